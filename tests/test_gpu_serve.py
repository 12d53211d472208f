"""Serving engine: continuous batching of independent SIMPLE cases.

The cases in flight are stepped together on one batch mesh; a case joins
while others run and leaves after its own last iteration, and the batch is
rebuilt each time with the continuing cases' state moved on the device.
Stated tolerance: every request's fields are bit-identical whether it is
served alone (one case in flight) or in the mix (several in flight, joining
and leaving at different iterations); the engine admits requests by HBM
budget and by equal settings."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERS = 25


def _case(x, ratio, cells, iters=ITERS, rans=True):
    import bench
    return bench.step_case(x.api, x, ratio, cells, iters, rans=rans)


def _submit(srv, x, case):
    import bench
    return bench.submit_case(srv, x.api, *case)


def _alone(x, cases, iters=ITERS):
    """Each case served on its own (a batch of one), one after another."""
    srv = x.hip.Server(0, slots=1, max_iterations=iters)
    out = []
    for case in cases:
        rid = _submit(srv, x, case)
        assert srv.wait(rid, 300.0)
        st = srv.status(rid)
        assert st["state"] == x._abi.REQ_DONE, st
        out.append((srv.fetch(rid), st))
    srv.close()
    return out


def _same(got, ref):
    assert set(got) == set(ref)
    for f in ref:
        assert np.array_equal(got[f], ref[f]), f


def test_serve_batch_alone_equals_mix(x):
    """Three k-omega steps (900, 1600 and 4500 cells) in flight together in
    one batch equal each served alone, bit for bit."""
    cases = [_case(x, r, c) for r, c in ((0.35, 900), (0.45, 1600), (0.5, 4500))]
    alone = _alone(x, cases)
    assert set(alone[0][0]) == {"U", "p", "k", "omega", "nut"}
    srv = x.hip.Server(0, slots=3, max_iterations=ITERS)
    ids = [_submit(srv, x, c) for c in cases]
    assert srv.wait(0, 300.0)
    spans = []
    for rid, (ref, st0) in zip(ids, alone):
        st = srv.status(rid)
        assert st["state"] == x._abi.REQ_DONE and st["iterations"] == st0["iterations"] == ITERS
        got = srv.fetch(rid)
        _same(got, ref)
        assert np.isfinite(got["U"]).all() and np.abs(got["U"]).max() > 1.0
        spans.append((st["t_start"], st["t_end"]))
    # stepped together: every case was in flight while every other was
    assert max(a for a, _ in spans) < min(b for _, b in spans), spans
    srv.close()


def test_serve_case_joins_running_batch(x):
    """A 4.5k-cell case is submitted while a 250k-cell case has run a few
    iterations: it joins the running batch (rebuilt with the large case's
    state moved on the device), and after the large case leaves it goes on
    alone (rebuilt again); both equal their alone runs bit for bit."""
    big = _case(x, 0.4, 250000, iters=12)
    small = _case(x, 0.55, 4500, iters=12)
    alone = _alone(x, [big, small], iters=12)
    srv = x.hip.Server(0, slots=4, max_iterations=12)
    b = _submit(srv, x, big)
    t0 = time.time()
    while srv.status(b)["iterations"] < 3:
        assert time.time() - t0 < 300.0
        time.sleep(0.01)
    s = _submit(srv, x, small)
    assert srv.wait(0, 300.0)
    sb, ss = srv.status(b), srv.status(s)
    assert sb["cells"] >= 240000 and ss["cells"] >= 4000
    assert ss["t_start"] < sb["t_end"]  # the small case ran inside the big one's run
    assert ss["batch_rebuilds"] >= 3     # big alone, big + small, big after small left
    _same(srv.fetch(b), alone[0][0])
    _same(srv.fetch(s), alone[1][0])
    srv.close()


def test_serve_state_moves_bit_for_bit(x, monkeypatch):
    """With XCAL_SERVE_REBUILD=1 the server rebuilds the batch after every
    iteration, moving every case's fields to a new batch mesh each time; the
    results still equal the alone runs bit for bit."""
    cases = [_case(x, 0.4, 6000, iters=8), _case(x, 0.55, 4500, iters=8)]
    alone = _alone(x, cases, iters=8)
    monkeypatch.setenv("XCAL_SERVE_REBUILD", "1")
    srv = x.hip.Server(0, slots=2, max_iterations=8)
    monkeypatch.delenv("XCAL_SERVE_REBUILD")
    ids = [_submit(srv, x, c) for c in cases]
    assert srv.wait(0, 300.0)
    assert srv.status(ids[0])["batch_rebuilds"] >= 8
    for rid, (ref, _) in zip(ids, alone):
        assert srv.status(rid)["iterations"] == 8
        _same(srv.fetch(rid), ref)
    srv.close()


def test_serve_keeps_unlike_cases_apart(x):
    """A laminar case between two k-omega cases waits for the batch to drain
    (settings differ), and every case still equals its alone run."""
    cases = [_case(x, 0.35, 900), _case(x, 0.5, 2500, rans=False), _case(x, 0.45, 1600)]
    alone = _alone(x, cases)
    assert set(alone[1][0]) == {"U", "p"}
    srv = x.hip.Server(0, slots=3, max_iterations=ITERS)
    ids = [_submit(srv, x, c) for c in cases]
    assert srv.wait(0, 300.0)
    st = [srv.status(i) for i in ids]
    assert st[1]["t_start"] >= st[0]["t_end"] and st[2]["t_start"] >= st[1]["t_end"]
    for rid, (ref, _) in zip(ids, alone):
        _same(srv.fetch(rid), ref)
    srv.close()


def test_serve_admission_respects_budget(x):
    """With an HBM budget that holds one case, three copies of a case run one
    after another even with three slots: no two are in flight together."""
    H = x.hip
    case = _case(x, 0.35, 900)
    srv = H.Server(0, slots=3, max_iterations=5)
    rid = _submit(srv, x, case)
    assert srv.wait(rid, 120.0)
    fp = srv.status(rid)["footprint"]
    assert fp > 0
    srv.close()
    srv = H.Server(0, slots=3, max_iterations=5, hbm_budget=int(fp * 1.5))
    ids = [_submit(srv, x, case) for _ in range(3)]
    assert srv.wait(0, 120.0)
    spans = sorted((srv.status(i)["t_start"], srv.status(i)["t_end"]) for i in ids)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1, spans
    srv.close()


def test_serve_rejects_oversized_case(x):
    H = x.hip
    case = _case(x, 0.5, 2500, rans=False)
    srv = H.Server(0, slots=1, max_iterations=5, hbm_budget=1 << 16)
    with pytest.raises(Exception, match="HBM budget"):
        _submit(srv, x, case)
    srv.close()


def test_serve_lanes_group_cases_by_size(x):
    """With two lanes, cases join the lane closest to them in size: the two
    60k-cell cases share one lane, the two 4.5k-cell cases the other, the
    lanes step concurrently, and every case equals its alone run bit for
    bit."""
    cases = [_case(x, 0.4, 60000, iters=10), _case(x, 0.55, 4500, iters=10),
             _case(x, 0.35, 56000, iters=10), _case(x, 0.5, 4800, iters=10)]
    alone = _alone(x, cases, iters=10)
    srv = x.hip.Server(0, slots=4, max_iterations=10, lanes=2)
    ids = [_submit(srv, x, c) for c in cases]
    assert srv.wait(0, 300.0)
    st = [srv.status(i) for i in ids]
    lanes = [s["lane"] for s in st]
    assert lanes[0] == lanes[2] and lanes[1] == lanes[3] and lanes[0] != lanes[1], lanes
    # the lanes were in flight at the same time
    span = {l: (min(s["t_start"] for s in st if s["lane"] == l),
                max(s["t_end"] for s in st if s["lane"] == l)) for l in set(lanes)}
    (a0, a1), (b0, b1) = span.values()
    assert a0 < b1 and b0 < a1, span
    for rid, (ref, _) in zip(ids, alone):
        assert srv.status(rid)["iterations"] == 10
        _same(srv.fetch(rid), ref)
    srv.close()
