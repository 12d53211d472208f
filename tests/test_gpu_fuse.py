"""Fused Krylov finalizes on small meshes.

With a diagonal preconditioner and at most 256 workgroups per launch, the
one-workgroup finalize after each reduction is folded into the next kernel
of the iteration: every workgroup reduces the partials in the same order and
applies the same scalar update.  The arithmetic is the same as with separate
finalizes (XCAL_NO_FUSE=1 at mesh creation), so SIMPLE runs match bit for
bit: CG for p and BiCGSTAB for U, FP64 and FP32 inner solves, and the k-omega
step's BiCGSTAB solves."""
import numpy as np
import pytest

from conftest import cavity_case
from test_gpu_komega import komega_bfs_case

pytestmark = pytest.mark.gpu


def _run(x, mesh, setup, model, iters, monkeypatch, fuse, turb=None):
    H = x.hip
    ctx = x.api._context(0)
    if not fuse:
        monkeypatch.setenv("XCAL_NO_FUSE", "1")
    dm = H.DeviceMesh(ctx, mesh)
    monkeypatch.delenv("XCAL_NO_FUSE", raising=False)
    sim = H.Simple(dm, setup)
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    if turb is not None:
        sim.set_turbulence(turb)
        t = model.turbulence
        for name in ("k", "omega", "nut"):
            sim.set_scalar(name, getattr(t, name))
    out = sim.run(iters)
    fields = [sim.get("U"), sim.get("p")]
    if turb is not None:
        fields += [sim.get("k"), sim.get("omega")]
    return out, fields


@pytest.mark.parametrize("prec", ["fp64", "fp32-ir"])
def test_fused_finalizes_bit_identical_cavity(x, meshes, monkeypatch, prec):
    api = x.api
    mesh = meshes("cavity40")
    model, config = cavity_case(api, mesh, iterations=40, precon="Jacobi", rtol_p=1e-3,
                                precision_p=prec)
    setup = api.simple_setup(model, config, pref=0.0)
    a, fa = _run(x, mesh, setup, model, 40, monkeypatch, True)
    b, fb = _run(x, mesh, setup, model, 40, monkeypatch, False)
    assert a["done"] == b["done"]
    assert np.array_equal(a["krylov_iters"], b["krylov_iters"])
    assert np.array_equal(a["residuals"], b["residuals"])
    for u, v in zip(fa, fb):
        assert np.array_equal(u, v)


def test_fused_finalizes_bit_identical_komega(x, meshes, monkeypatch):
    api = x.api
    mesh = meshes("bfs10")
    model, config = komega_bfs_case(api, mesh, 30, "Jacobi")
    setup = api.simple_setup(model, config)
    ts = api.turbulence_setup(model, config)
    a, fa = _run(x, mesh, setup, model, 30, monkeypatch, True, ts)
    b, fb = _run(x, mesh, setup, model, 30, monkeypatch, False, ts)
    assert a["done"] == b["done"]
    assert np.array_equal(a["krylov_iters"], b["krylov_iters"])
    assert np.array_equal(a["residuals"], b["residuals"])
    for u, v in zip(fa, fb):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("prec", ["fp64", "fp32-ir"])
def test_fused_batch_chains_bit_identical(x, monkeypatch, prec):
    """Batched CG and BiCGSTAB with Jacobi: each kernel of an iteration does
    the previous kernel's per-case finalize itself (2 and 3 launches per
    iteration).  Three cavities of different size and Reynolds number run as
    one batch match the separate-finalize chain (XCAL_NO_FUSE=1) bit for
    bit: fields, residuals and Krylov iteration counts per case."""
    api, H = x.api, x.hip
    ms = [x.mesh.cavity2d(n) for n in (12, 20, 16)]
    model, config = cavity_case(api, ms[0], iterations=25, precon="Jacobi", rtol_p=1e-3,
                                precision_p=prec)
    setup = api.simple_setup(model, config, pref=0.0)
    ctx = x.api._context(0)
    runs = []
    for fuse in (True, False):
        if not fuse:
            monkeypatch.setenv("XCAL_NO_FUSE", "1")
        dm = H.DeviceMesh.batch(ctx, ms)
        monkeypatch.delenv("XCAL_NO_FUSE", raising=False)
        sim = H.Simple(dm, setup)
        sim.set_case_nu([0.01, 0.004, 0.0025])
        out = sim.run_batch(25)
        runs.append((out, sim.get("U"), sim.get("p")))
    (a, Ua, pa), (b, Ub, pb) = runs
    assert a["done"] == b["done"] == 25
    assert np.array_equal(a["krylov_iters"], b["krylov_iters"])
    assert np.array_equal(a["residuals"], b["residuals"])
    assert np.array_equal(Ua, Ub) and np.array_equal(pa, pb)
