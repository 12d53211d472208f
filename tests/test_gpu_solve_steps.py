"""The steps of solve_system! around the Krylov solve on the device
(Solve_1_api.jl:268-307): the Jacobi smoother (Smoothers_jacobi.jl:54-110)
and Crank-Nicolson's x = 2x - x_old (:291-296), plus the run! options that
reach them (inner_loops, the turbulence model's share of the stop test).

The smoother keeps the reference's row order and is bit-exact against the
oracle.  Whole runs differ from the oracle only by the order of the Krylov
reductions; stated tolerances beside each test."""
import numpy as np
import pytest

from conftest import (cavity_case, komega_bfs_case, patch_face_average, rel, steady_bfs_case)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


def _system(x, A, O, ctx, m, rng):
    """A convection-diffusion matrix on the device and in the oracle."""
    H = x.hip
    n, nf = m.n_cells, m.n_faces
    dm = H.DeviceMesh(ctx, m)
    om = O.OracleMesh(m, A)
    mdh, gh = rng.standard_normal(nf) * m.face_area, 0.5 + rng.random(nf)
    md = H.Field(dm, A.FACE_SCALAR).upload(mdh)
    g = H.Field(dm, A.FACE_SCALAR).upload(gh)
    src = rng.standard_normal(n)
    s = H.Field(dm, A.CELL_SCALAR).upload(src)
    eq = H.Equation(dm, False)
    eq.discretise([(A.TIME_STEADY, 1, 1.0), (A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)],
                  [(s, 1)])
    nz, b = om.discretise_scalar([(A.TIME_STEADY, 1, 1.0), (A.DIV_UPWIND, 1, mdh),
                                  (A.LAPLACIAN_LINEAR, -1, gh)], [(src, 1)])
    return dm, om, eq, nz, b


@pytest.mark.parametrize("name", ["cavity16", "box12", "trig40"])
@pytest.mark.parametrize("loops,omega", [(1, 1.0), (5, 2.0 / 3.0)])
def test_jacobi_smoother_bit_exact(x, A, O, ctx, meshes, name, loops, omega):
    """apply_smoother! on the device equals the oracle bit for bit.  The
    Krylov solve after it is given rtol 1, so it stops before its first
    iteration and leaves the smoothed field."""
    H = x.hip
    m = meshes(name)
    rng = np.random.default_rng(11)
    dm, om, eq, nz, b = _system(x, A, O, ctx, m, rng)
    x0 = rng.standard_normal(m.n_cells)
    phi = H.Field(dm, A.CELL_SCALAR).upload(x0)
    eq.precond_update(A.PRECON_JACOBI)
    setup = A.SolverSetupC(A.SOLVER_BICGSTAB, A.PRECON_JACOBI, A.PREC_FP64, 100, 1e-7, 1.0, 0.0,
                           1.0, 0.0, loops, omega)
    niter, _, _ = eq.solve(setup, phi)
    assert niter == 0
    assert np.array_equal(phi.download(), om.smooth_jacobi(nz, b, x0, loops, omega))


def test_crank_nicolson_solve_step(x, A, O, ctx, meshes):
    """An equation whose first term is Time{CrankNicolson} takes
    x = 2 x* - x_old after its solve (x* the Krylov solution, x_old the
    smoothed start), against the dense solution within 1e-10."""
    H = x.hip
    m = meshes("cavity16")
    rng = np.random.default_rng(12)
    dm = H.DeviceMesh(ctx, m)
    om = O.OracleMesh(m, A)
    prev = rng.standard_normal(m.n_cells)
    pf = H.Field(dm, A.CELL_SCALAR).upload(prev)
    x0 = rng.standard_normal(m.n_cells)
    out = {}
    for tt in (A.TIME_EULER, A.TIME_CRANK_NICOLSON):
        eq = H.Equation(dm, False)
        eq.discretise([(tt, 1, 1.0), (A.LAPLACIAN_LINEAR, -1, 0.1)], [], pf, dt=0.01)
        eq.precond_update(A.PRECON_DILU)
        phi = H.Field(dm, A.CELL_SCALAR).upload(x0)
        setup = A.SolverSetupC(A.SOLVER_CG, A.PRECON_DILU, A.PREC_FP64, 1000, 1e-7, 1.0, 1e-300,
                               1e-14, 0.0, 2, 2.0 / 3.0)
        eq.solve(setup, phi)
        out[tt] = phi.download()
    nz, b = om.discretise_scalar([(A.TIME_EULER, 1, 1.0), (A.LAPLACIAN_LINEAR, -1, 0.1)], [], prev,
                                 dt=0.01)
    rp, cv = om.csr()
    M = np.zeros((m.n_cells, m.n_cells))
    for i in range(m.n_cells):
        M[i, cv[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    xs = np.linalg.solve(M, b)
    xsm = om.smooth_jacobi(nz, b, x0, 2, 2.0 / 3.0)
    scale = np.abs(xs).max()
    assert np.allclose(out[A.TIME_EULER], xs, rtol=0, atol=1e-10 * scale)
    assert np.allclose(out[A.TIME_CRANK_NICOLSON], 2 * xs - xsm, rtol=0, atol=1e-10 * scale)


def test_piso_crank_nicolson_matches_oracle(x, A, O, meshes):
    """40x40 cavity from rest, 20 PISO steps with a Time{CrankNicolson}
    momentum term (dt 0.005, tight solves): U and p within 1e-8 of the
    oracle, and different from the Euler run."""
    api = x.api
    mesh = meshes("cavity40")
    out = {}
    for scheme in (api.Euler, api.CrankNicolson):
        model, config = cavity_case(api, mesh, iterations=20, rtol_p=1e-10)
        model.time = api.Transient()
        config.schemes["U"] = api.Schemes(divergence=api.Upwind, time=scheme)
        config.runtime.time_step = 0.005
        config.solvers["U"].rtol = 1e-10
        config.solvers["U"].relax = 1.0
        setup = api.simple_setup(model, config, pref=0.0)
        ref = O.OracleMesh(mesh, A).piso_run(setup, model.momentum.U, model.momentum.p, 20, 2)
        api.run(model, config, pref=0.0)
        eU, ep = rel(model.momentum.U, ref["U"]), rel(model.momentum.p, ref["p"])
        print(f"PISO {scheme.__name__}: U {eU:.2e} p {ep:.2e}")
        assert eU < 1e-8 and ep < 1e-8
        out[scheme] = model.momentum.U.copy()
    assert rel(out[api.CrankNicolson], out[api.Euler]) > 1e-6


def test_piso_inner_loops_reach_the_device(x, A, O, meshes):
    """run!(model, config; inner_loops=3) takes three pressure correctors per
    time step, as the oracle's PISO with three (within 1e-8), and differs
    from the default two."""
    api = x.api
    mesh = meshes("cavity40")
    res = {}
    for inner in (2, 3):
        model, config = cavity_case(api, mesh, iterations=10, rtol_p=1e-10)
        model.time = api.Transient()
        config.schemes["U"] = api.Schemes(divergence=api.Upwind, time=api.Euler)
        config.runtime.time_step = 0.02
        config.solvers["U"].rtol = 1e-10
        config.solvers["U"].relax = 1.0
        setup = api.simple_setup(model, config, pref=0.0)
        ref = O.OracleMesh(mesh, A).piso_run(setup, model.momentum.U, model.momentum.p, 10, inner)
        api.run(model, config, pref=0.0, inner_loops=inner)
        assert rel(model.momentum.U, ref["U"]) < 1e-8 and rel(model.momentum.p, ref["p"]) < 1e-8
        res[inner] = model.momentum.p.copy()
    assert rel(res[3], res[2]) > 1e-8
    with pytest.raises(ValueError):
        api.run(model, config, inner_loops=0)


def test_steady_bfs_with_smoother_reference_check(x, A, O, meshes):
    """The reference's steady laminar BFS test (JacobiSmoother on U and p,
    2d_incompressible_laminar_BFS.jl) on the device: fields within 1e-6 of
    the oracle after 500 iterations, and the outlet's average velocity half
    the inlet's within 0.1 U."""
    api = x.api
    mesh = meshes("bfs10")
    model, config = steady_bfs_case(api, mesh)
    setup = api.simple_setup(model, config)
    ref = O.OracleMesh(mesh, A).simple_run(setup, model.momentum.U, model.momentum.p, 500)
    config.runtime.write_interval = -1
    res = api.run(model, config)
    assert len(res["p"]) == ref["done"]
    eU, ep = rel(model.momentum.U, ref["U"]), rel(model.momentum.p, ref["p"])
    print(f"steady BFS with smoother: U {eU:.2e} p {ep:.2e}")
    assert eU < 1e-6 and ep < 1e-6
    outlet = patch_face_average(mesh, model.momentum.U, "outlet")
    assert np.linalg.norm(outlet - np.array([0.25, 0.0, 0.0])) <= 0.05, outlet


def test_komega_write_interval_keeps_the_device_stop(x, A, meshes, tmp_path):
    """With k and omega held to a tighter convergence than U and p, a run
    that writes every 7 iterations stops where the unwritten run stops (the
    turbulence residuals take part in the stop test), with the same
    residual history and fields."""
    api = x.api
    mesh = meshes("bfs10")
    out = []
    for wi in (-1, 7):
        model, config = komega_bfs_case(api, mesh, 60, "DILU")
        for f in ("U", "p"):
            config.solvers[f].convergence = 1e-2
        for f in ("k", "omega"):
            config.solvers[f].convergence = 1e-12
        config.runtime.write_interval = wi
        res = api.run(model, config, directory=str(tmp_path))
        out.append((res, model.momentum.U.copy(), model.turbulence.k.copy()))
    (ra, Ua, ka), (rb, Ub, kb) = out
    assert len(ra["p"]) == len(rb["p"]) == 60
    assert np.array_equal(ra["p"], rb["p"]) and np.array_equal(Ua, Ub) and np.array_equal(ka, kb)
