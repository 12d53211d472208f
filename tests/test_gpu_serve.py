"""Serving engine: continuous admission of independent SIMPLE cases.

Stated tolerance: every request's fields are bit-identical whether it runs
alone (one slot, requests one at a time), in the mix (several slots, all
requests in flight at once) or through run!() on its own; the engine admits
requests by HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERS = 25


def _requests(x):
    import bench
    api = x.api
    return [bench.step_case(api, x, ratio, cells, ITERS, rans=rans)
            for ratio, cells, rans in ((0.35, 900, True), (0.5, 2500, False), (0.45, 1600, True))]


def _submit(srv, x, case):
    import bench
    return bench.submit_case(srv, x.api, *case)


def test_serve_alone_equals_mix(x):
    H = x.hip
    reqs = _requests(x)
    alone = []
    srv = H.Server(0, slots=1, max_iterations=ITERS)
    for case in reqs:
        rid = _submit(srv, x, case)
        assert srv.wait(rid, 120.0)
        assert srv.status(rid)["iterations"] == ITERS
        alone.append(srv.fetch(rid))
    srv.close()
    assert set(alone[0]) == {"U", "p", "k", "omega", "nut"} and set(alone[1]) == {"U", "p"}
    srv = H.Server(0, slots=3, max_iterations=ITERS)
    ids = [_submit(srv, x, case) for case in reqs]
    assert srv.wait(0, 120.0)
    for rid, ref in zip(ids, alone):
        st = srv.status(rid)
        assert st["state"] == x._abi.REQ_DONE and st["iterations"] == ITERS
        got = srv.fetch(rid)
        for f in ref:
            assert np.array_equal(got[f], ref[f]), f
        assert np.isfinite(got["U"]).all() and np.abs(got["U"]).max() > 1.0
    srv.close()
    # the same k-omega case through run!() on the caller's own context
    mesh, model, config = reqs[0]
    x.api.run(model, config)
    for f, v in (("U", model.momentum.U), ("p", model.momentum.p), ("k", model.turbulence.k),
                 ("omega", model.turbulence.omega), ("nut", model.turbulence.nut)):
        assert np.array_equal(np.asarray(v).reshape(alone[0][f].shape), alone[0][f]), f


def test_serve_admission_respects_budget(x):
    """With an HBM budget that holds one case, three copies of a case run one
    after another even with three slots: no two are in flight together."""
    H = x.hip
    case = _requests(x)[0]
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
    case = _requests(x)[1]
    srv = H.Server(0, slots=1, max_iterations=5, hbm_budget=1 << 16)
    with pytest.raises(Exception, match="HBM budget"):
        _submit(srv, x, case)
    srv.close()
