"""k-omega RANS step (RANS_kOmega.jl) on the device against the oracle.

Case: the reference's examples/2D_BFS_highRe.jl on its backwardFacingStep_10mm
mesh (k and omega wall functions, Jacobi preconditioners).  The wall functions
use log() whose last-ulp rounding differs between the device and host
libraries, and Krylov reductions sum in a different order, so the parity is
tolerance-based.  Stated tolerance: after 100 SIMPLE iterations U, p, k,
omega and nut within 1e-6 relative L2 of the oracle."""
import numpy as np
import pytest

from conftest import komega_bfs_case, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precon", ["Jacobi", "DILU"])
def test_komega_bfs_matches_oracle(x, A, O, meshes, precon):
    api = x.api
    mesh = meshes("bfs10")
    iters = 100
    model, config = komega_bfs_case(api, mesh, iters, precon)
    setup = api.simple_setup(model, config)
    ts = api.turbulence_setup(model, config)
    t = model.turbulence
    ref = O.OracleMesh(mesh, A).simple_run(setup, model.momentum.U, model.momentum.p, iters,
                                           turb=ts, k=t.k, omega=t.omega, nut=t.nut)
    res = api.run(model, config)
    assert len(res["p"]) == ref["done"]
    errs = dict(U=rel(model.momentum.U, ref["U"]), p=rel(model.momentum.p, ref["p"]),
                k=rel(t.k, ref["k"]), omega=rel(t.omega, ref["omega"]), nut=rel(t.nut, ref["nut"]))
    print(f"k-omega BFS ({precon}): " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert all(v < 1e-6 for v in errs.values()), errs
    assert np.all(t.k > 0) and np.all(t.omega > 0)


def test_komega_serving_size_step_within_rounding_spread(x, A, O):
    """A backward-facing step of the serving workload's larger kind (about
    100k cells, k-omega at U 10, nu 1e-5, DILU), 2 SIMPLE iterations from
    the uniform inlet state.  This start is chaotic at the rounding level:
    the oracle itself, rerun with its initial U multiplied by
    (1 + 1e-13 N(0,1)), moves by 3e-6 to 8e-6 in U, k and omega depending on
    the seed (1e-15 gives 3e-6), so a fixed 1e-6 bound cannot hold for any
    two summation orders.  Stated tolerance: each of U, p, k, omega and nut
    within 4x that oracle spread (the largest over seeds 0, 1, 2; relative
    L2) of the oracle, and below 1e-4 outright; every pressure solve of the
    oracle converges."""
    import bench
    api = x.api
    mesh, model, config = bench.step_case(api, x, 0.45, 1e5, 2, precon="DILU")
    setup = api.simple_setup(model, config)
    ts = api.turbulence_setup(model, config)
    t = model.turbulence
    om = O.OracleMesh(mesh, A)
    ref = om.simple_run(setup, model.momentum.U, model.momentum.p, 2,
                        turb=ts, k=t.k, omega=t.omega, nut=t.nut)
    fields = ("U", "p", "k", "omega", "nut")
    spread = dict.fromkeys(fields, 0.0)
    for seed in range(3):
        rng = np.random.default_rng(seed)
        U2 = model.momentum.U * (1 + 1e-13 * rng.standard_normal(model.momentum.U.shape))
        ref2 = om.simple_run(setup, U2, model.momentum.p, 2, turb=ts, k=t.k, omega=t.omega,
                             nut=t.nut)
        spread = {f: max(spread[f], rel(ref2[f], ref[f])) for f in fields}
    hist = api.run(model, config)
    assert np.all(ref["krylov_iters"][3] < 1000)
    dev = dict(U=model.momentum.U, p=model.momentum.p, k=t.k, omega=t.omega, nut=t.nut)
    errs = {f: rel(dev[f], ref[f]) for f in fields}
    print(f"k-omega step {mesh.n_cells} cells: " +
          " ".join(f"{f} {errs[f]:.2e} (spread {spread[f]:.2e})" for f in fields))
    for f in fields:
        assert errs[f] < 4 * max(spread[f], 1e-9) and errs[f] < 1e-4, (f, errs[f], spread[f])
    assert len(hist["p"]) == 2
