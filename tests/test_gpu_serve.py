"""Serving engine: continuous admission of independent SIMPLE cases.

Stated tolerance: every request's U and p are bit-identical whether it runs
alone (one slot, requests one at a time) or in the mix (several slots, all
requests in flight at once); the engine admits requests by HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(x):
    import bench
    api = x.api
    out = []
    for ratio, cells in ((0.35, 900), (0.5, 2500), (0.45, 1600)):
        mesh, model, config = bench.step_case(api, x, ratio, cells, 25)
        out.append((mesh, model, api.simple_setup(model, config)))
    return out


def test_serve_alone_equals_mix(x):
    H = x.hip
    reqs = _requests(x)
    alone = []
    srv = H.Server(0, slots=1, max_iterations=25)
    for mesh, model, setup in reqs:
        rid = srv.submit(mesh, setup, model.momentum.U, model.momentum.p)
        assert srv.wait(rid, 120.0)
        assert srv.status(rid)["iterations"] == 25
        alone.append(srv.fetch(rid))
    srv.close()
    srv = H.Server(0, slots=3, max_iterations=25)
    ids = [srv.submit(mesh, setup, model.momentum.U, model.momentum.p) for mesh, model, setup in reqs]
    assert srv.wait(0, 120.0)
    for rid, (U1, p1) in zip(ids, alone):
        st = srv.status(rid)
        assert st["state"] == x._abi.REQ_DONE and st["iterations"] == 25
        U, p = srv.fetch(rid)
        assert np.array_equal(U, U1) and np.array_equal(p, p1)
        assert np.isfinite(U).all() and np.abs(U).max() > 1.0
    srv.close()


def test_serve_admission_respects_budget(x):
    """With an HBM budget that holds one case, three copies of a case run one
    after another even with three slots: no two are in flight together."""
    H = x.hip
    mesh, model, setup = _requests(x)[0]
    srv = H.Server(0, slots=3, max_iterations=5)
    rid = srv.submit(mesh, setup, model.momentum.U, model.momentum.p)
    assert srv.wait(rid, 120.0)
    fp = srv.status(rid)["footprint"]
    assert fp > 0
    srv.close()
    srv = H.Server(0, slots=3, max_iterations=5, hbm_budget=int(fp * 1.5))
    ids = [srv.submit(mesh, setup, model.momentum.U, model.momentum.p) for _ in range(3)]
    assert srv.wait(0, 120.0)
    spans = sorted((srv.status(i)["t_start"], srv.status(i)["t_end"]) for i in ids)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1, spans
    srv.close()
