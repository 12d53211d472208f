"""Transient PISO (Solvers_2_PISO.jl) and the Time{Euler} term on the device
against the oracle.

Assembly with Time{Euler} keeps the reference's evaluation order and is
bit-exact.  Whole PISO runs differ from the oracle only by the order of the
Krylov reductions; stated tolerances: 1e-8 relative L2 with tight solves
(rtol 1e-10), 1e-6 for the reference's own BFS settings (rtol 1e-3) after
100 steps."""
import numpy as np
import pytest

from conftest import cavity_case, patch_face_average, rel, transient_bfs_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


@pytest.mark.parametrize("name", ["cavity40", "box12", "step"])
def test_euler_discretise_bit_exact(x, A, O, ctx, meshes, name):
    """discretise! with Time{Euler} and runtime.dt (xcal_discretise_dt), scalar
    and vector equations, against the oracle bit for bit."""
    m = meshes(name)
    H = x.hip
    rng = np.random.default_rng(3)
    n, nf = m.n_cells, m.n_faces
    dm = H.DeviceMesh(ctx, m)
    om = O.OracleMesh(m, A)
    mdh, gh = rng.standard_normal(nf) * m.face_area, 0.5 + rng.random(nf)
    md = H.Field(dm, A.FACE_SCALAR).upload(mdh)
    g = H.Field(dm, A.FACE_SCALAR).upload(gh)
    dt = 0.0137
    prev_s = rng.standard_normal(n)
    prev_v = rng.standard_normal((3, n))
    src_s = rng.standard_normal(n)
    ps = H.Field(dm, A.CELL_SCALAR).upload(prev_s)
    pv = H.Field(dm, A.CELL_VECTOR).upload(prev_v)
    ss = H.Field(dm, A.CELL_SCALAR).upload(src_s)
    terms = [(A.TIME_EULER, 1, 1.0), (A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)]
    oterms = [(A.TIME_EULER, 1, 1.0), (A.DIV_UPWIND, 1, mdh), (A.LAPLACIAN_LINEAR, -1, gh)]
    eq = H.Equation(dm, False)
    eq.discretise(terms, [(ss, 1)], ps, dt=dt)
    onz, ob = om.discretise_scalar(oterms, [(src_s, 1)], prev_s, dt)
    nz, b = eq.download()
    assert np.array_equal(nz, onz) and np.array_equal(b, ob)
    ev = H.Equation(dm, True)
    ev.discretise(terms, [], pv, dt=dt)
    onz, ob = om.discretise_vector(oterms, [], prev_v, dt)
    for c in range(3):
        nz0, b = ev.download(which=1, component=c)  # A0 and b of component c
        assert np.array_equal(nz0, onz) and np.array_equal(b, ob[c])


def _tight_transient_cavity(api, mesh, steps):
    model, config = cavity_case(api, mesh, iterations=steps, rtol_p=1e-10)
    model.time = api.Transient()
    config.schemes["U"] = api.Schemes(divergence=api.Upwind, time=api.Euler)
    config.runtime.time_step = 0.02
    config.solvers["U"].rtol = 1e-10
    config.solvers["U"].relax = 1.0
    return model, config


def test_piso_cavity_matches_oracle(x, A, O, meshes):
    """40x40 lid-driven cavity started from rest, 30 PISO steps of dt 0.02
    with 2 correctors; fields within 1e-8 of the oracle."""
    api = x.api
    mesh = meshes("cavity40")
    model, config = _tight_transient_cavity(api, mesh, 30)
    setup = api.simple_setup(model, config, pref=0.0)
    ref = O.OracleMesh(mesh, A).piso_run(setup, model.momentum.U, model.momentum.p, 30, 2)
    res = api.run(model, config, pref=0.0)
    assert len(res["p"]) == ref["done"] == 30
    eU, ep = rel(model.momentum.U, ref["U"]), rel(model.momentum.p, ref["p"])
    print(f"PISO cavity40: U {eU:.2e} p {ep:.2e}")
    assert eU < 1e-8 and ep < 1e-8


def test_piso_single_corrector_equals_simple_on_device(x, A, meshes):
    """On the device too, one PISO step with one corrector is one SIMPLE
    iteration with the same Time{Euler} momentum equation and an unrelaxed
    pressure: bit for bit."""
    api = x.api
    H = x.hip
    mesh = meshes("cavity40")
    model, config = _tight_transient_cavity(api, mesh, 10)
    config.solvers["p"].relax = 1.0
    for s in config.solvers.values():
        s.convergence = 0.0
    setup = api.simple_setup(model, config, pref=0.0)
    ctx = api._context(0)
    out = []
    for piso in (False, True):
        sim = H.Simple(H.DeviceMesh(ctx, mesh), setup)
        sim.set_U(model.momentum.U)
        sim.set_p(model.momentum.p)
        r = sim.run_piso(10, 1) if piso else sim.run(10)
        out.append((r, sim.get("U"), sim.get("p")))
    (ra, Ua, pa), (rb, Ub, pb) = out
    assert ra["done"] == rb["done"] == 10
    assert np.array_equal(ra["residuals"], rb["residuals"])
    assert np.array_equal(Ua, Ub) and np.array_equal(pa, pb)


def test_piso_transient_bfs_reference_case(x, A, O, meshes):
    """The reference's transient laminar BFS test on its mesh: after 100
    steps the device fields are within 1e-6 of the oracle's; after 1000
    the outlet's average velocity is half the inlet's within 0.1 U
    (2d_incompressible_transient_laminar_BFS.jl)."""
    api = x.api
    mesh = meshes("bfs10")
    model, config = transient_bfs_case(api, mesh, 100)
    setup = api.simple_setup(model, config)
    ref = O.OracleMesh(mesh, A).piso_run(setup, model.momentum.U, model.momentum.p, 100, 2)
    api.run(model, config)
    eU, ep = rel(model.momentum.U, ref["U"]), rel(model.momentum.p, ref["p"])
    print(f"PISO BFS 100 steps: U {eU:.2e} p {ep:.2e}")
    assert eU < 1e-6 and ep < 1e-6
    model, config = transient_bfs_case(api, mesh, 1000)
    res = api.run(model, config)
    assert len(res["p"]) == 1000
    outlet = patch_face_average(mesh, model.momentum.U, "outlet")
    assert np.linalg.norm(outlet - np.array([0.25, 0.0, 0.0])) <= 0.05, outlet


def test_piso_nonorth_correctors_match_oracle(x, A, O, meshes):
    """PISO with one non-orthogonal corrector per pressure corrector
    (Solvers_2_PISO.jl:148-165) on the jittered box channel as a transient
    case: 5 steps, U and p within 1e-8 of the oracle."""
    import bench
    api = x.api
    mesh = meshes("box12")
    model, config = bench.transient_channel(api, mesh, 5, 0.05)
    setup = api.simple_setup(model, config, ncorrectors=1)
    ref = O.OracleMesh(mesh, A).piso_run(setup, model.momentum.U, model.momentum.p, 5, 2)
    api.run(model, config, ncorrectors=1)
    eU, ep = rel(model.momentum.U, ref["U"]), rel(model.momentum.p, ref["p"])
    print(f"PISO ncorrectors=1: U {eU:.2e} p {ep:.2e}")
    assert eU < 1e-8 and ep < 1e-8
