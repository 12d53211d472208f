"""Pipelined one-workgroup DILU sweeps on single meshes.

A run of small DILU levels is swept by one workgroup that prefetches rows
several levels ahead and keeps recent outputs in an LDS ring (preloaded with
the rows before the run).  It keeps the level-stepping kernels' row-to-thread
assignment and summation orders, so SIMPLE runs (CG and BiCGSTAB with DILU,
FP64 and FP32 inner solves) are bit-identical either way."""
import numpy as np
import pytest

from conftest import cavity_case, channel_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


def _pair(x, ctx, mesh, monkeypatch):
    H = x.hip
    dm = H.DeviceMesh(ctx, mesh)
    monkeypatch.setenv("XCAL_NO_PIPE", "1")
    ref = H.DeviceMesh(ctx, mesh)
    monkeypatch.delenv("XCAL_NO_PIPE")
    assert ref.sweep_form() == (-1, -1)
    return dm, ref


def _simple(x, dm, model, config, iters):
    api = x.api
    sim = x.hip.Simple(dm, api.simple_setup(model, config))
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    out = sim.run(iters)
    return sim.get("U"), sim.get("p"), out["residuals"], out["krylov_iters"]


@pytest.mark.parametrize("name,case,prec", [
    ("cavity40", "cavity", "fp64"), ("step", "cavity", "fp64"), ("step60k", "cavity", "fp64"),
    ("box40", "channel", "fp64"), ("box40", "channel", "fp32-ir")])
def test_pipelined_runs_equal_level_stepping(x, ctx, meshes, monkeypatch, name, case, prec):
    """box40: a naturally numbered 40^3 box, whose middle DILU levels get
    launches of their own and whose end levels form runs that depend on them
    (the ring preload).  step60k (two mesher blocks) has dependencies further
    back than the ring, which its run gathers from global memory."""
    api = x.api
    if name == "box40":
        mesh = x.mesh.box3d(40, 40, 40, jitter=0.1, seed_perm=0)
    elif name == "step60k":
        ny = 110
        mesh = x.mesh.backward_facing_step_2d(0.45, 1.0, 1.0, 4.0, nx_in=ny, nx_out=4 * ny,
                                              ny_low=50, ny_up=60)
    else:
        mesh = meshes(name)
    dm, ref = _pair(x, ctx, mesh, monkeypatch)
    assert dm.sweep_form()[0] > 0 and dm.sweep_form()[1] > 0
    iters = 12
    runs = []
    for d in (dm, ref):
        if case == "cavity":
            if name.startswith("step"):
                model, config = _step_on(api, mesh, iters)
            else:
                model, config = cavity_case(api, mesh, iterations=iters)
        else:
            model, config = channel_case(api, mesh, iterations=iters, rtol_p=1e-8, precision_p=prec)
        runs.append(_simple(x, d, model, config, iters))
    (U0, p0, r0, k0), (U1, p1, r1, k1) = runs
    assert np.array_equal(U0, U1) and np.array_equal(p0, p1)
    assert np.array_equal(r0, r1) and np.array_equal(k0, k1)
    assert np.isfinite(U0).all() and np.abs(U0).max() > 0.1


def _step_on(api, mesh, iters):
    """Laminar flow over the test step mesh (patches inlet, outlet, walls, top)."""
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=0.01))
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [1.0, 0, 0]), api.Extrapolated("outlet"),
                        api.Wall("walls", [0, 0, 0]), api.Wall("top", [0, 0, 0])],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Extrapolated("walls"), api.Extrapolated("top")])
    S = lambda s, r: api.SolverSetup(solver=s, preconditioner=api.DILU(), convergence=1e-7,
                                     relax=r)
    config = api.Configuration(
        solvers=dict(U=S(api.Bicgstab(), 0.7), p=S(api.Cg(), 0.3)),
        schemes=dict(U=api.Schemes(divergence=api.Upwind), p=api.Schemes()),
        runtime=api.Runtime(iterations=iters, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config
