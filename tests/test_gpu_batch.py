"""Batched form: independent cases in one batch mesh, stepped together.

A case's results must not depend on which other cases share the batch: each
case keeps its own Krylov iterations, workgroups never straddle two cases and
each case's reductions are summed in a fixed order over its own workgroups.
So a case run in a batch of several must equal, bit for bit, the same case
run as a batch of one.  Against the CPU oracle the batched solves are held
to the tolerances of the single-mesh solver.
"""
import numpy as np
import pytest

from conftest import cavity_case, rel

pytestmark = pytest.mark.gpu

SIZES = (12, 20, 16)


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


def _cases(x):
    return [x.mesh.cavity2d(n) for n in SIZES]


def _case_fields(mesh, seed):
    rng = np.random.default_rng(seed)
    return dict(gamma=0.5 + rng.random(mesh.n_faces),
                mdot=rng.standard_normal(mesh.n_faces) * mesh.face_area * 0.2,
                src=rng.standard_normal(mesh.n_cells), phi=0.01 * rng.standard_normal(mesh.n_cells))


def _system(x, A, dm, meshes, sym):
    """Assemble Laplacian (+ upwind when not symmetric) on a (batch) mesh whose
    cases are `meshes`; returns (equation, phi field, per-case fields)."""
    H = x.hip
    F = [_case_fields(m, 100 + SIZES.index(int(round(np.sqrt(m.n_cells))))) for m in meshes]
    cat = lambda key: np.concatenate([f[key] for f in F])
    g = H.Field(dm, A.FACE_SCALAR).upload(cat("gamma"))
    md = H.Field(dm, A.FACE_SCALAR).upload(np.zeros(dm.mesh.n_faces) if sym else cat("mdot"))
    terms = [(A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)]
    src = H.Field(dm, A.CELL_SCALAR).upload(cat("src"))
    phi = H.Field(dm, A.CELL_SCALAR).upload(cat("phi"))
    eq = H.Equation(dm, False)
    eq.discretise(terms, [(src, 1)], phi)
    bcs = A.bc_array([(A.BC_DIRICHLET, i, 1.0) for i in range(len(meshes[0].patch_names))])
    eq.apply_bcs(terms, bcs, phi)
    return eq, phi, F, terms, bcs


def test_batch_mesh_is_block_diagonal(x, ctx):
    ms = _cases(x)
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    assert dm.ncase == 3 and dm.n == sum(m.n_cells for m in ms)
    rp, cv = dm.csr_pattern()
    off = 0
    for m in ms:
        r1, c1 = x.hip.DeviceMesh(ctx, m).csr_pattern()
        assert np.array_equal(rp[off:off + m.n_cells + 1] - rp[off], r1)
        assert np.array_equal(cv[rp[off]:rp[off + m.n_cells]] - off, c1)
        off += m.n_cells


def test_batch_assembly_equals_single_meshes(x, A, ctx):
    """Assembly kernels are per cell and per face: the batch's nzval and b are
    the cases' single-mesh results concatenated, bit for bit."""
    ms = _cases(x)
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    eq, _, _, _, _ = _system(x, A, dm, ms, sym=False)
    nzb, bb = eq.download()
    zo = co = 0
    for m in ms:
        d1 = x.hip.DeviceMesh(ctx, m)
        e1, _, _, _, _ = _system(x, A, d1, [m], sym=False)
        nz1, b1 = e1.download()
        assert np.array_equal(nzb[zo:zo + len(nz1)], nz1)
        assert np.array_equal(bb[co:co + m.n_cells], b1)
        zo += len(nz1)
        co += m.n_cells


@pytest.mark.parametrize("solver,precon,prec", [
    ("SOLVER_CG", "PRECON_DILU", "PREC_FP64"), ("SOLVER_CG", "PRECON_JACOBI", "PREC_FP64"),
    ("SOLVER_BICGSTAB", "PRECON_DILU", "PREC_FP64"), ("SOLVER_BICGSTAB", "PRECON_JACOBI", "PREC_FP64"),
    ("SOLVER_CG", "PRECON_DILU", "PREC_FP32_IR")])
def test_batch_krylov_alone_equals_mix_and_oracle(x, A, O, ctx, solver, precon, prec):
    """Stated tolerances: bit-identical alone vs in the batch; against the
    oracle's FP64 solve (rtol 1e-12) relative L2 <= 1e-9 (FP64) or 1e-8
    (FP32 inner solves in FP64 refinement to 1e-12)."""
    ms = _cases(x)
    sym = solver == "SOLVER_CG"
    pre = getattr(A, precon)
    setup = A.SolverSetupC(getattr(A, solver), pre, getattr(A, prec), 4000, 1e-7, 1.0, 0.0, 1e-12,
                           1e-4)
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    eq, phi, F, _, bcs = _system(x, A, dm, ms, sym)
    eq.precond_update(pre)
    its, outer, res, rc = eq.solve_batch(setup, phi)
    assert rc == 0
    sol = phi.download()
    off = 0
    for k, m in enumerate(ms):
        d1 = x.hip.DeviceMesh.batch(ctx, [m])
        e1, p1, _, _, _ = _system(x, A, d1, [m], sym)
        e1.precond_update(pre)
        it1, out1, res1, rc1 = e1.solve_batch(setup, p1)
        assert np.array_equal(sol[off:off + m.n_cells], p1.download()), f"case {k}"
        assert its[k] == it1[0] and outer[k] == out1[0] and res[k] == res1[0]
        om = O.OracleMesh(m, A)
        f = F[k]
        oterms = [(A.DIV_UPWIND, 1, np.zeros(m.n_faces) if sym else f["mdot"]),
                  (A.LAPLACIAN_LINEAR, -1, f["gamma"])]
        onz, ob = om.discretise_scalar(oterms, [(f["src"], 1)])
        om.apply_bcs(oterms, bcs, -1, f["phi"], onz, ob)
        ox, oit, orc = om.krylov(getattr(A, solver), pre, onz, ob, f["phi"], itmax=4000,
                                 atol=1e-30, rtol=1e-12)
        e = rel(sol[off:off + m.n_cells], ox)
        print(f"case {k} n={m.n_cells}: its {its[k]} (oracle {oit}), sweeps {outer[k]}, rel {e:.2e}")
        assert e < (1e-9 if prec == "PREC_FP64" else 1e-8)
        off += m.n_cells


def _batch_setup(api, meshes, iterations, precon="DILU"):
    model, config = cavity_case(api, meshes[0], iterations=iterations, precon=precon)
    return api.simple_setup(model, config, pref=0.0)


def test_batch_simple_alone_equals_mix_and_oracle(x, A, O, ctx):
    """Three lid-driven cavities of different size and Reynolds number run as
    one batch for 40 SIMPLE iterations.  Stated tolerances: each case
    bit-identical to the same case run alone (a batch of one); U and p within
    1e-6 relative L2 of the oracle's run of that case."""
    api = x.api
    ms = _cases(x)
    nus = [0.01, 0.004, 0.0025]
    iters = 40
    setup = _batch_setup(api, ms, iters)
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    sim = x.hip.Simple(dm, setup)
    sim.set_case_nu(nus)
    out = sim.run_batch(iters)
    assert out["done"] == iters
    U, p = sim.get("U"), sim.get("p")
    off = 0
    for k, m in enumerate(ms):
        d1 = x.hip.DeviceMesh.batch(ctx, [m])
        s1 = x.hip.Simple(d1, setup)
        s1.set_case_nu([nus[k]])
        o1 = s1.run_batch(iters)
        sl = slice(off, off + m.n_cells)
        assert np.array_equal(U[:, sl], s1.get("U")), f"case {k} U"
        assert np.array_equal(p[sl], s1.get("p")), f"case {k} p"
        assert np.array_equal(out["residuals"][k], o1["residuals"][0])
        model, config = cavity_case(api, m, nu=nus[k], iterations=iters)
        ref = O.OracleMesh(m, A).simple_run(api.simple_setup(model, config, pref=0.0),
                                             model.momentum.U, model.momentum.p, iters)
        eU, ep = rel(U[:, sl], ref["U"]), rel(p[sl], ref["p"])
        print(f"case {k} n={m.n_cells} nu={nus[k]}: rel L2 U {eU:.2e} p {ep:.2e}")
        assert eU < 1e-6 and ep < 1e-6
        off += m.n_cells


@pytest.mark.parametrize("prec", ["PREC_FP64", "PREC_FP32_IR"])
def test_batch_pipelined_sweeps_equal_level_stepping(x, A, ctx, monkeypatch, prec):
    """The pipelined per-case DILU sweeps (register prefetch, LDS ring) keep
    the level-stepping kernels' row-to-thread assignment and summation order:
    CG and BiCGSTAB solves and a SIMPLE run are bit-identical either way."""
    ms = _cases(x)
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    assert dm.sweep_form() == (2, 2)  # quads: at most 2 lower / upper entries
    monkeypatch.setenv("XCAL_NO_PIPE", "1")
    dm_ref = x.hip.DeviceMesh.batch(ctx, ms)
    monkeypatch.delenv("XCAL_NO_PIPE")
    assert dm_ref.sweep_form() == (-1, -1)
    for solver in ("SOLVER_CG", "SOLVER_BICGSTAB"):
        sym = solver == "SOLVER_CG"
        setup = A.SolverSetupC(getattr(A, solver), A.PRECON_DILU, getattr(A, prec), 4000, 1e-7,
                               1.0, 0.0, 1e-12, 1e-4)
        out = []
        for d in (dm, dm_ref):
            eq, phi, _, _, _ = _system(x, A, d, ms, sym)
            eq.precond_update(A.PRECON_DILU)
            its, outer, res, rc = eq.solve_batch(setup, phi)
            assert rc == 0
            out.append((phi.download(), list(its), list(res)))
        assert np.array_equal(out[0][0], out[1][0]), solver
        assert out[0][1] == out[1][1] and out[0][2] == out[1][2], solver
    api = x.api
    setup = _batch_setup(api, ms, 15)
    runs = []
    for d in (dm, dm_ref):
        sim = x.hip.Simple(d, setup)
        sim.set_case_nu([0.01, 0.004, 0.0025])
        sim.run_batch(15)
        runs.append((sim.get("U"), sim.get("p")))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_batch_far_dependencies_equal_level_stepping(x, A, ctx, monkeypatch):
    """A 2D step mesh of about 60k cells has DILU dependencies more than 4096
    rows back; the pipelined batch sweeps read those from the global copy of
    the forward sweep's output, which each sweep must write afresh.  Solves on
    such a batch equal the level-stepping kernels bit for bit, twice in a row
    on the same equation (the second sweep finds the first one's values in the
    work vectors)."""
    ms = [x.mesh.backward_facing_step_2d(0.4, 1.0, 1.0, 4.0, nx_in=113, nx_out=452,
                                         ny_low=45, ny_up=68),
          x.mesh.backward_facing_step_2d(0.5, 1.0, 1.0, 4.0, nx_in=20, nx_out=80,
                                         ny_low=10, ny_up=10)]
    dm = x.hip.DeviceMesh.batch(ctx, ms)
    assert dm.n > 60000
    assert dm.sweep_form()[0] > 0 and dm.sweep_far()[0]
    monkeypatch.setenv("XCAL_NO_PIPE", "1")
    dm_ref = x.hip.DeviceMesh.batch(ctx, ms)
    monkeypatch.delenv("XCAL_NO_PIPE")
    H = x.hip
    rng = np.random.default_rng(7)
    nf = dm.mesh.n_faces
    gamma = 0.5 + rng.random(nf)
    mdot = rng.standard_normal(nf) * np.concatenate([m.face_area for m in ms]) * 0.2
    src = rng.standard_normal(dm.n)
    for solver in ("SOLVER_CG", "SOLVER_BICGSTAB"):
        sym = solver == "SOLVER_CG"
        setup = A.SolverSetupC(getattr(A, solver), A.PRECON_DILU, A.PREC_FP64, 60, 1e-7,
                               1.0, 0.0, 1e-12, 1e-4)
        out = []
        for d in (dm, dm_ref):
            g = H.Field(d, A.FACE_SCALAR).upload(gamma)
            md = H.Field(d, A.FACE_SCALAR).upload(np.zeros(nf) if sym else mdot)
            terms = [(A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)]
            s = H.Field(d, A.CELL_SCALAR).upload(src)
            phi = H.Field(d, A.CELL_SCALAR).upload(np.zeros(d.n))
            eq = H.Equation(d, False)
            eq.discretise(terms, [(s, 1)], phi)
            bcs = A.bc_array([(A.BC_DIRICHLET, i, 1.0) for i in range(len(ms[0].patch_names))])
            eq.apply_bcs(terms, bcs, phi)
            eq.precond_update(A.PRECON_DILU)
            runs = []
            for _ in range(2):
                phi.upload(np.zeros(d.n))
                its, _, res, rc = eq.solve_batch(setup, phi)
                assert rc in (0, 1)  # 1: a case stopped at 60 iterations
                runs.append((phi.download(), list(its), list(res)))
            assert np.array_equal(runs[0][0], runs[1][0]), solver
            out.append(runs[0])
        assert np.array_equal(out[0][0], out[1][0]), solver
        assert out[0][1] == out[1][1] and out[0][2] == out[1][2], solver


@pytest.mark.parametrize("prec", ["PREC_FP64", "PREC_FP32_IR"])
def test_batch_pipe_rows_per_thread_bit_identical(x, A, ctx, monkeypatch, prec):
    """The pipelined batch sweeps at 1, 2 and 4 rows of a level per thread
    (256, 128 and 64 threads per case) give the same CG and BiCGSTAB solves
    bit for bit, dot products included; the tuner picks one of them and
    reports a timing for each."""
    ms = _cases(x)
    out = {}
    for rows in ("1", "2", "4", None):
        if rows:
            monkeypatch.setenv("XCAL_PIPE_ROWS", rows)
        else:
            monkeypatch.delenv("XCAL_PIPE_ROWS", raising=False)
        dm = x.hip.DeviceMesh.batch(ctx, ms)
        res = []
        for solver in ("SOLVER_CG", "SOLVER_BICGSTAB"):
            sym = solver == "SOLVER_CG"
            setup = A.SolverSetupC(getattr(A, solver), A.PRECON_DILU, getattr(A, prec), 4000, 1e-7,
                                   1.0, 0.0, 1e-12, 1e-4)
            eq, phi, _, _, _ = _system(x, A, dm, ms, sym)
            eq.precond_update(A.PRECON_DILU)
            its, outer, r, rc = eq.solve_batch(setup, phi)
            assert rc == 0
            res.append((phi.download(), list(its), list(r)))
        f, b, t = dm.pipe_tuning()
        if rows:
            assert (f, b) == (int(rows), int(rows))
        else:
            assert f in (1, 2, 4) and b in (1, 2, 4)
            assert all(v > 0 for v in t.values()), t
            print("tuned rows per thread", f, b, t)
        out[rows] = res
    for rows in ("2", "4", None):
        for (u, iu, ru), (v, iv, rv) in zip(out["1"], out[rows]):
            assert np.array_equal(u, v) and iu == iv and ru == rv, rows


def test_batch_graph_replay_equals_eager(x, A, ctx, monkeypatch):
    """Batched Krylov iterations replayed from HIP graphs (two iterations per
    graph, captured on first use and replayed on every later solve, with
    assembly and host reads of the stop flags between replays) give the same
    SIMPLE run bit for bit as eager launches (the default), for DILU and
    Jacobi preconditioning."""
    ms = _cases(x)
    api = x.api
    for precon in ("DILU", "Jacobi"):
        runs = []
        for graph in ("0", "1"):
            monkeypatch.setenv("XCAL_GRAPH", graph)
            dm = x.hip.DeviceMesh.batch(ctx, ms)
            setup = _batch_setup(api, ms, 12, precon=precon)
            sim = x.hip.Simple(dm, setup)
            sim.set_case_nu([0.01, 0.004, 0.0025])
            out = sim.run_batch(12)
            runs.append((sim.get("U"), sim.get("p"), out))
        monkeypatch.delenv("XCAL_GRAPH")
        assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
