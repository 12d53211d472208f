"""Pinning the CPU oracle against the reference's golden vectors and against
what the mathematics fixes (CPU only).  A plausible slip in the restatement
(a dropped boundary term, a wrong sign or index, a transposed operand) fails
at least one of these."""
import json

import numpy as np
import pytest

from conftest import (GOLDEN, bfs_verification_profile, cavity_case, channel_case,
                      docs_bfs_case, patch_face_average, rel, steady_bfs_case,
                      transient_bfs_case)


def dense(om, nz):
    rp, cv = om.csr()
    M = np.zeros((om.n, om.n))
    for i in range(om.n):
        M[i, cv[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    return M


def test_laplace_golden_A_b(A, O, meshes):
    """unit_test_laplace.jl:115-130 on laplace_unit_3by3.unv (k=1, rDf=1)."""
    g = json.load(open(GOLDEN / "laplace_3by3.json"))
    m = meshes("laplace3")
    om = O.OracleMesh(m, A)
    idx = {n: i for i, n in enumerate(m.patch_names)}
    bcs = A.bc_array([(A.BC_DIRICHLET, idx["left_wall"], 50.0),
                      (A.BC_ZEROGRADIENT, idx["right_wall"], 0.0),
                      (A.BC_DIRICHLET, idx["bottom_wall"], 10.0),
                      (A.BC_ZEROGRADIENT, idx["upper_wall"], 0.0)])
    T = np.full(m.n_cells, 15.0)
    terms = [(A.TIME_STEADY, 1, 0.0), (A.LAPLACIAN_LINEAR, -1, 1.0)]
    nz, b = om.discretise_scalar(terms, [(np.zeros(m.n_cells), -1)], T)
    om.apply_bcs(terms, bcs, -1, T, nz, b)
    assert np.allclose(dense(om, nz), np.array(g["A"], float), atol=g["A_atol"])
    assert np.allclose(dense(om, nz), np.array(g["A"], float), atol=1e-12)
    assert np.allclose(b, g["b"], rtol=1e-12)


@pytest.mark.parametrize("name", ["cavity40", "box12", "trig40", "cavity_hex"])
def test_csr_pattern_closed_form(A, O, meshes, name):
    """row i = {i} U neighbours(i), ascending (sparse_matrix_connectivity)."""
    m = meshes(name)
    rp, cv = O.OracleMesh(m, A).csr()
    own = m.face_owner.reshape(-1, 2)[m.n_bfaces:] - 1
    rows = [{i} for i in range(m.n_cells)]
    for a, b in own:
        rows[a].add(b)
        rows[b].add(a)
    for i in range(m.n_cells):
        assert list(cv[rp[i]:rp[i + 1]]) == sorted(rows[i])


def test_uniform_laplacian_and_upwind_closed_form(A, O, meshes):
    """On the uniform 40x40 cavity: Laplacian off-diagonals are -Gamma*A/delta
    = -Gamma, diagonals the negated row sum; upwind rows sum to sum of outflow
    minus inflow of the face fluxes."""
    m = meshes("cavity40")
    om = O.OracleMesh(m, A)
    nz, _ = om.discretise_scalar([(A.LAPLACIAN_LINEAR, -1, 0.01)], [])
    D = dense(om, nz)
    off = D - np.diag(np.diag(D))
    assert np.allclose(off[off != 0], -0.01, rtol=1e-12)
    assert np.allclose(np.diag(D), -off.sum(1), rtol=1e-12)
    rng = np.random.default_rng(0)
    mdot = rng.standard_normal(m.n_faces)
    nz, _ = om.discretise_scalar([(A.DIV_UPWIND, 1, mdot)], [])
    D = dense(om, nz)
    net = np.zeros(m.n_cells)
    cf = m.cell_faces - 1
    rng_ = m.cell_faces_range.reshape(-1, 2)
    cell_of = np.repeat(np.arange(m.n_cells), rng_[:, 1] - rng_[:, 0] + 1)
    np.add.at(net, cell_of, mdot[cf] * m.cell_nsign)
    assert np.allclose(D.sum(1), net, atol=1e-12)           # upwind row sum = net outflow
    assert np.all(np.diag(D) >= 0) and np.all(D - np.diag(np.diag(D)) <= 0)


@pytest.mark.parametrize("name", ["cavity40", "box12_natural", "trig40", "cavity_hex"])
def test_green_gauss_exact_for_linear_fields(A, O, meshes, name):
    """With exact face-centroid values of phi = a.x + c, Green-Gauss returns a
    in every closed cell with planar faces; the tensor variant returns the
    outer product for a linear vector field.  (Jittered hexes have non-planar
    faces, for which the identity holds only to O(jitter).)"""
    m = meshes(name)
    om = O.OracleMesh(m, A)
    a = np.array([0.3, -1.7, 0.9 if m.dim == 3 else 0.0])
    fc = m.face_centre.reshape(-1, 3)
    phif = fc @ a + 2.5
    g = np.stack(om.green_gauss(phif))
    tol = 1e-9 if name in ("cavity_hex", "trig40") else 1e-10
    assert np.allclose(g, a[:, None], atol=tol * np.abs(a).max() * 10)
    B = np.array([[1.0, 2.0, 0.5], [-0.5, 0.25, 1.0], [0.0, 3.0, -1.0]])
    if m.dim == 2:
        B[:, 2] = 0.0
    Uf = [np.ascontiguousarray(fc @ B[p]) for p in range(3)]
    T = np.stack(om.green_gauss_tensor(*Uf)).reshape(3, 3, -1)
    assert np.allclose(T, B[:, :, None], atol=1e-8)


@pytest.mark.parametrize("name", ["cavity40", "box12", "step"])
def test_interpolation_and_divergence(A, O, meshes, name):
    """Uniform flow has zero divergence; linear interpolation reproduces a
    field that is linear along the owner-neighbour line at the weight point."""
    m = meshes(name)
    om = O.OracleMesh(m, A)
    U = np.array([0.7, -0.2, 0.4 if m.dim == 3 else 0.0])
    ones = np.ones(m.n_faces)
    md = om.flux(U[0] * ones, U[1] * ones, U[2] * ones)
    assert np.allclose(om.div(md), 0.0, atol=1e-10)
    cc = m.cell_centre.reshape(-1, 3)
    phi = cc @ U
    pf = om.interpolate(phi)
    own = m.face_owner.reshape(-1, 2) - 1
    w = m.face_weight
    expect = w * phi[own[:, 0]] + (1 - w) * phi[own[:, 1]]
    assert np.allclose(pf, expect, rtol=1e-14, atol=1e-14)
    if name == "cavity40":   # uniform: weight 1/2 at the face midpoint -> exact
        fc = m.face_centre.reshape(-1, 3)
        assert np.allclose(pf[m.n_bfaces:], (fc @ U)[m.n_bfaces:], atol=1e-13)


def _random_mesh_matrix(A, O, m, seed=0, asym=True):
    """Diagonally dominant matrix with the mesh's sparsity pattern."""
    om = O.OracleMesh(m, A)
    rng = np.random.default_rng(seed)
    rp, cv = om.csr()
    nz = -rng.random(len(cv)) * (1.0 if asym else 0.0)
    if not asym:
        own = m.face_owner.reshape(-1, 2)[m.n_bfaces:] - 1
        for a, b in own:
            v = -rng.random()
            nz[rp[a] + np.searchsorted(cv[rp[a]:rp[a + 1]], b)] = v
            nz[rp[b] + np.searchsorted(cv[rp[b]:rp[b + 1]], a)] = v
    for i in range(om.n):
        k = rp[i] + np.searchsorted(cv[rp[i]:rp[i + 1]], i)
        nz[k] = 0.0
        nz[k] = -nz[rp[i]:rp[i + 1]].sum() + 0.75 + rng.random()
    return om, nz


@pytest.mark.parametrize("name", ["cavity16", "box12"])
def test_dilu_identities(A, O, meshes, name):
    """test_DILU.jl:24-40: diag((D*+L) D*^-1 (D*+U)) = diag(A); forward and
    backward substitution equal dense triangular solves."""
    m = meshes(name)
    om, nz = _random_mesh_matrix(A, O, m, seed=1)
    Ad = dense(om, nz)
    D = om.dilu_diagonal(nz)
    L = np.tril(Ad, -1)
    U = np.triu(Ad, 1)
    PL = (np.diag(D) + L) @ np.diag(1.0 / D)
    PU = np.diag(D) + U
    P = PL @ PU
    assert np.allclose(np.diag(P), np.diag(Ad), rtol=1e-12)
    b = P @ np.ones(om.n)
    x = np.linalg.solve(PU, np.linalg.solve(PL, b))
    assert np.allclose(om.dilu_apply(nz, D, b), x, rtol=1e-10)
    assert np.allclose(x, 1.0, rtol=1e-8)


@pytest.mark.parametrize("solver,precon,asym", [(0, 2, False), (0, 1, False), (0, 0, False),
                                               (1, 2, True), (1, 1, True), (1, 0, True)])
def test_krylov_against_dense_solve(A, O, meshes, solver, precon, asym):
    m = meshes("box12")
    om, nz = _random_mesh_matrix(A, O, m, seed=2, asym=asym)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(om.n)
    xs = np.linalg.solve(dense(om, nz), b)
    x0 = rng.standard_normal(om.n) * 0.1
    x, it, rc = om.krylov(solver, precon, nz, b, x0, itmax=1000, atol=1e-14, rtol=1e-13)
    assert rc == 0 and 0 < it < 1000
    assert rel(x, xs) < 1e-10
    assert om.residual(nz, b, x) < 1e-11


def test_krylov_itmax_flag(A, O, meshes):
    m = meshes("box12")
    om, nz = _random_mesh_matrix(A, O, m, seed=4, asym=False)
    b = np.ones(om.n)
    _, it, rc = om.krylov(0, 1, nz, b, np.zeros(om.n), itmax=2, atol=0.0, rtol=1e-14)
    assert it == 2 and rc == 1


def test_residual_definition(A, O, meshes):
    """residual = sqrt(sum (b-Ax)^2)/||b|| with denominator 1 when b = 0."""
    m = meshes("cavity16")
    om, nz = _random_mesh_matrix(A, O, m, seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(om.n)
    b = rng.standard_normal(om.n)
    r = b - dense(om, nz) @ x
    assert np.isclose(om.residual(nz, b, x), np.linalg.norm(r) / np.linalg.norm(b), rtol=1e-12)
    z = np.zeros(om.n)
    assert np.isclose(om.residual(nz, z, x), np.linalg.norm(dense(om, nz) @ x), rtol=1e-12)


def test_implicit_relaxation_and_reference(A, O, meshes):
    m = meshes("cavity16")
    om, nz = _random_mesh_matrix(A, O, m, seed=7)
    Ad = dense(om, nz)
    rng = np.random.default_rng(8)
    b = rng.standard_normal(om.n)
    phi = rng.standard_normal(om.n)
    nz2, b2 = nz.copy(), b.copy()
    om.relax_implicit_diagdom(nz2, b2, phi, 0.7)
    d0 = np.diag(Ad)
    dm = np.maximum(np.abs(d0), np.abs(Ad).sum(1) - np.abs(d0)) / 0.7
    assert np.allclose(np.diag(dense(om, nz2)), dm, rtol=1e-14)
    assert np.allclose(b2, b + (dm - d0) * phi, rtol=1e-12)
    nz3, b3 = nz.copy(), b.copy()
    om.set_reference(nz3, b3, 2.0, 0)
    assert b3[0] == d0[0] * 2.0 and dense(om, nz3)[0, 0] == 2 * d0[0]


def test_cavity_simple_physics(x, A, O, meshes):
    """40x40 lid-driven cavity at Re=100: residuals fall and the u-velocity on
    the vertical centreline reaches the Ghia et al. (1982) minimum of about
    -0.21 near y = 0.45 (first-order upwind on 40x40 gives about -0.19)."""
    api = x.api
    m = meshes("cavity40")
    model, config = cavity_case(api, m, iterations=400)
    setup = api.simple_setup(model, config, pref=0.0)
    out = O.OracleMesh(m, A).simple_run(setup, model.momentum.U, model.momentum.p, 400)
    r = out["residuals"]
    assert r[0, -1] < 1e-3 * r[0, 0] and r[3, -1] < 1e-2 * r[3, 0]
    cc = m.cell_centre.reshape(-1, 3)
    col = np.abs(cc[:, 0] - 0.5 - 1 / 80) < 1e-9
    u = out["U"][0][col]
    assert -0.24 < u.min() < -0.15
    assert 0.35 < cc[col][np.argmin(u), 1] < 0.55


def test_channel_simple_mass_conservation(x, A, O, meshes):
    """Developing laminar flow in the jittered box: the outlet mass flux
    approaches the inlet's as SIMPLE converges (the corrected flux uses the
    under-relaxed pressure, so the balance closes only at convergence) and
    the velocity stays bounded."""
    api = x.api
    m = meshes("box12")
    om = O.OracleMesh(m, A)
    pr = m.patch_range.reshape(-1, 2)
    imbalance = []
    for its in (30, 150):
        model, config = channel_case(api, m, iterations=its)
        out = om.simple_run(api.simple_setup(model, config), model.momentum.U,
                            model.momentum.p, its)
        md = out["mdotf"]
        inflow = md[pr[0, 0] - 1:pr[0, 1]].sum()
        outflow = md[pr[1, 0] - 1:pr[1, 1]].sum()
        assert np.isclose(inflow, -1.0, rtol=1e-12)          # Dirichlet U=1 on a unit inlet
        imbalance.append(abs(outflow + inflow))
        assert np.abs(out["U"]).max() < 3.0
    assert imbalance[1] < 2e-3 and imbalance[1] < imbalance[0] / 3


@pytest.mark.parametrize("name", ["step", "box12"])
def test_empty_bc_closed_form(A, O, meshes, name):
    """Empty gives no coefficient from any operator (empty.jl:46-106): its
    assembly equals Zerogradient's once the flux through its faces is zero
    (Zerogradient keeps the convective a_p = F).  Its face values are zero
    (empty_interpolation.jl)."""
    m = meshes(name)
    om = O.OracleMesh(m, A)
    rng = np.random.default_rng(11)
    nf = m.n_faces
    gamma = 0.5 + rng.random(nf)
    mdot = rng.standard_normal(nf) * m.face_area
    phi = rng.standard_normal(m.n_cells)
    last = len(m.patch_names) - 1
    pr = np.asarray(m.arrays["patch_range"]).reshape(-1, 2) - 1 if "patch_range" in m.arrays else None
    lo, hi = int(pr[last][0]), int(pr[last][1])
    base = [(A.BC_DIRICHLET, i, 1.0 + i) for i in range(last)]
    empty = A.bc_array(base + [(A.BC_EMPTY, last, 0.0)])
    zg = A.bc_array(base + [(A.BC_ZEROGRADIENT, last, 0.0)])
    md0 = mdot.copy()
    md0[lo:hi + 1] = 0.0
    for bcs, md in ((empty, mdot), (zg, md0)):
        terms = [(A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, gamma)]
        nz, b = om.discretise_scalar(terms, [], phi)
        om.apply_bcs(terms, bcs, -1, phi, nz, b)
        if bcs is empty:
            nz_e, b_e = nz, b
        else:
            # the interior faces of the two runs see different fluxes only on
            # boundary faces, which interior coefficients never read
            assert np.array_equal(nz_e, nz) and np.array_equal(b_e, b)
    pf = om.interpolate(phi)
    om.correct_boundaries(empty, phi, pf)
    assert np.all(pf[lo:hi + 1] == 0.0)


def test_midpoint_gradient_exact_for_linear_fields(A, O, meshes):
    """Grad{Midpoint} (Calculate_0_gradient.jl:235-244): on a uniform mesh the
    face midpoint of the two cell centres is the face centre, so a linear
    field's face values, and hence its Green-Gauss gradient in cells whose
    faces are all internal, are exact; the corrected face values equal the
    field at the face centres."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    a = np.array([0.7, -1.3, 0.0])
    cc = m.cell_centre.reshape(-1, 3)
    phi = cc @ a
    bcs = A.bc_array([(A.BC_EXTRAPOLATED, i, 0.0) for i in range(len(m.patch_names))])
    f, g = om.grad_midpoint(bcs, phi)
    fc = m.face_centre.reshape(-1, 3)
    nbf = m.n_bfaces
    assert np.allclose(f[0][nbf:], (fc @ a)[nbf:], rtol=0, atol=1e-13)
    interior = np.setdiff1d(np.arange(m.n_cells), m.boundary_cellsID - 1)
    assert np.allclose(g[0][interior], a[0], atol=1e-12)
    assert np.allclose(g[1][interior], a[1], atol=1e-12)
    # vector field: each component's gradient row
    U = [cc @ a, cc @ (2 * a), cc @ (-a)]
    fv, gv = om.grad_midpoint(bcs, U)
    for p, s in enumerate((1.0, 2.0, -1.0)):
        assert np.allclose(gv[3 * p][interior], s * a[0], atol=1e-12)
        assert np.allclose(gv[3 * p + 1][interior], s * a[1], atol=1e-12)


@pytest.mark.parametrize("name", ["cavity16", "box12"])
def test_euler_time_term_closed_form(A, O, meshes, name):
    """Time{Euler} (Discretise_1_schemes.jl:25-42) adds V/dt to the diagonal
    and prev V/dt to b, and nothing off the diagonal: the assembly with it
    minus the assembly with Time{SteadyState} is exactly that, per
    component of a vector equation."""
    m = meshes(name)
    om = O.OracleMesh(m, A)
    rng = np.random.default_rng(5)
    n, nf = m.n_cells, m.n_faces
    mdot, gam = rng.standard_normal(nf) * m.face_area, 0.5 + rng.random(nf)
    prev = rng.standard_normal((3, n))
    src = rng.standard_normal((3, n))
    dt = 0.0137
    out = {}
    for kind in (A.TIME_STEADY, A.TIME_EULER):
        terms = [(kind, 1, 1.0), (A.DIV_UPWIND, 1, mdot), (A.LAPLACIAN_LINEAR, -1, gam)]
        out[kind] = om.discretise_vector(terms, [([src[0], src[1], src[2]], -1)], prev, dt)
    (nzS, bS), (nzE, bE) = out[A.TIME_STEADY], out[A.TIME_EULER]
    rp, cv = om.csr()
    diag = np.array([rp[i] + np.searchsorted(cv[rp[i]:rp[i + 1]], i) for i in range(n)])
    off = np.ones(len(nzS), bool)
    off[diag] = False
    assert np.array_equal(nzS[off], nzE[off])
    V = m.cell_volume
    assert np.allclose(nzE[diag] - nzS[diag], V / dt, rtol=1e-12, atol=0)
    for c in range(3):
        assert np.allclose(bE[c] - bS[c], prev[c] * V / dt, rtol=1e-9, atol=1e-12 * np.abs(bE[c]).max())


def test_piso_single_corrector_is_simple_with_time_term(x, A, O, meshes):
    """One PISO time step with one corrector (Solvers_2_PISO.jl:107-175) is
    one SIMPLE iteration (Solvers_1_SIMPLE.jl:164-258) of the same
    Time{Euler} momentum equation with the pressure unrelaxed: the same
    calls in the same order.  So the two loops agree bit for bit, step by
    step, which pins the PISO loop to the SIMPLE loop the physical checks
    above pin."""
    api = x.api
    m = meshes("cavity16")
    model, config = cavity_case(api, m, iterations=12)
    config.schemes["U"] = api.Schemes(divergence=api.Upwind, time=api.Euler)
    config.runtime.time_step = 0.05
    config.solvers["p"].relax = 1.0
    for s in config.solvers.values():
        s.convergence = 0.0  # SIMPLE takes every iteration, as PISO does
    setup = api.simple_setup(model, config)
    om = O.OracleMesh(m, A)
    a = om.simple_run(setup, model.momentum.U, model.momentum.p, 12)
    b = om.piso_run(setup, model.momentum.U, model.momentum.p, 12, inner_loops=1)
    assert a["done"] == b["done"] == 12
    for k in ("U", "p", "mdotf", "residuals", "krylov_iters"):
        assert np.array_equal(a[k], b[k]), k
    c = om.piso_run(setup, model.momentum.U, model.momentum.p, 12, inner_loops=2)
    assert not np.array_equal(a["p"], c["p"])  # the second corrector does work


def test_piso_transient_bfs_reference_check(x, A, O, meshes):
    """The reference's transient laminar BFS test
    (2d_incompressible_transient_laminar_BFS.jl): after 1000 PISO steps the
    outlet's average velocity is half the inlet's, within 0.1 U."""
    api = x.api
    m = meshes("bfs10")
    model, config = transient_bfs_case(api, m, 1000)
    setup = api.simple_setup(model, config)
    out = O.OracleMesh(m, A).piso_run(setup, model.momentum.U, model.momentum.p, 1000, 2)
    assert out["done"] == 1000
    inlet = np.array([0.5, 0.0, 0.0])
    outlet = patch_face_average(m, out["U"], "outlet")
    assert np.linalg.norm(outlet - 0.5 * inlet) <= 0.1 * 0.5, outlet
    assert np.all(np.isfinite(out["U"])) and np.all(np.isfinite(out["p"]))


def test_nonorth_correction_exact_for_linear_fields(A, O, meshes):
    """nonorthogonal_face_correction (Solvers_1_SIMPLE.jl:267-336) with the
    exact gradient of a linear field a.x: the implicit Laplacian carries
    a.E_f per face and the correction a.T_hat, so together a.S_f, whose sum
    over a closed cell is zero.  On the jittered box the uncorrected
    residual of interior cells is O(1e-3) of the scale; corrected it is at
    rounding level."""
    m = meshes("box12")
    om = O.OracleMesh(m, A)
    n, nf = m.n_cells, m.n_faces
    a = np.array([0.7, -1.3, 0.4])
    phi = m.cell_centre.reshape(-1, 3) @ a
    grad = np.tile(a[:, None], (1, n))
    terms = [(A.LAPLACIAN_LINEAR, -1, np.ones(nf))]
    nz, b = om.discretise_scalar(terms, [(np.zeros(n), -1)], phi)
    Aphi = om.spmv(nz, phi)
    scale = np.abs(Aphi).max()
    interior = np.setdiff1d(np.arange(n), m.boundary_cellsID - 1)
    r0 = (Aphi - b)[interior]
    r1 = (Aphi - om.nonorth_correction(grad, phi, np.ones(nf), b))[interior]
    assert np.abs(r0).max() > 1e-2 * scale
    assert np.abs(r1).max() < 1e-12 * scale


def _chain5(A, O, meshes):
    m = meshes("chain5")
    om = O.OracleMesh(m, A)
    rp, cv = om.csr()
    assert [list(cv[rp[i]:rp[i + 1]]) for i in range(5)] == \
        [[0, 1], [0, 1, 2], [1, 2, 3], [2, 3, 4], [3, 4]]
    return om


def test_jacobi_smoother_golden(A, O, meshes):
    """The system of test_smoothers.jl:32-46 (300/200 diagonal, -100
    off-diagonals, b = [1e5, 0, 0, 0, 2e4], solution [460, 380, 300, 220,
    140]) with JacobiSmoother(loops, 1) from zero.  The reference asserts
    1e-10 after 50 loops, which no Jacobi iteration reaches on this matrix:
    its iteration matrix I - D^-1 A has spectral radius 0.816, so 50 loops
    leave an error of 2.5e-2 in norm.  Pinned here: the error after L loops
    is the iteration matrix's power applied to the start error (to rounding)
    for L = 50, and 200 loops reach the dense solution within 1e-10."""
    om = _chain5(A, O, meshes)
    v = np.array([300, -100, -100, 200, -100, -100, 200, -100, -100, 200, -100, -100, 300.0])
    b = np.array([200 * 500, 0, 0, 0, 200 * 100.0])
    M = dense(om, v)
    xs = np.linalg.solve(M, b)
    assert np.allclose(xs, [460, 380, 300, 220, 140], rtol=1e-14)
    G = np.eye(5) - np.linalg.solve(np.diag(np.diag(M)), M)
    rho = np.abs(np.linalg.eigvals(G)).max()
    assert 0.81 < rho < 0.82
    e50 = np.linalg.matrix_power(G, 50) @ (np.zeros(5) - xs)
    x50 = om.smooth_jacobi(v, b, np.zeros(5), 50, 1.0)
    assert np.allclose(x50 - xs, e50, rtol=1e-8, atol=1e-10)
    assert 1e-2 < np.linalg.norm(x50 - xs) < 5e-2
    x200 = om.smooth_jacobi(v, b, np.zeros(5), 200, 1.0)
    assert np.allclose(x200, xs, rtol=0, atol=1e-10)


def test_jacobi_smoother_weighted_sweeps(A, O, meshes):
    """Two sweeps of weight 2/3 equal the matrix form
    x <- omega D^-1 (b - (A - D) x) + (1 - omega) x, every row reading the
    previous sweep (a Gauss-Seidel order or a dropped weight differs)."""
    om = _chain5(A, O, meshes)
    rng = np.random.default_rng(5)
    v = np.array([300, -100, -100, 200, -100, -100, 200, -100, -100, 200, -100, -100, 300.0])
    v = v * rng.uniform(0.9, 1.1, v.size)
    M = dense(om, v)
    D = np.diag(np.diag(M))
    b = rng.standard_normal(5)
    x0 = rng.standard_normal(5)
    w = 2 / 3
    ref = x0.copy()
    for _ in range(2):
        ref = w * np.linalg.solve(D, b - (M - D) @ ref) + (1 - w) * ref
    assert np.allclose(om.smooth_jacobi(v, b, x0, 2, w), ref, rtol=1e-13)


def test_smoother_constructors(x):
    """JacobiSmoother(mesh): 5 loops, omega 1; JacobiSmoother(domain, loops):
    omega 2/3 (test_smoothers.jl:16-30)."""
    api = x.api
    m = x.mesh.cavity2d(4)
    s = api.JacobiSmoother(m)
    assert (s.loops, s.omega) == (5, 1.0)
    s = api.JacobiSmoother(domain=m, loops=5)
    assert s.loops == 5 and s.omega == pytest.approx(2 / 3)
    c = api.SolverSetup(solver=api.Cg(), preconditioner=api.Jacobi(), convergence=1e-7,
                        relax=1.0, smoother=s).c()
    assert c.smoother_loops == 5 and c.smoother_omega == pytest.approx(2 / 3)
    assert api.SolverSetup(solver=api.Cg(), preconditioner=api.Jacobi(), convergence=1e-7,
                           relax=1.0).c().smoother_loops == 0


def test_crank_nicolson_step_closed_form(A, O, meshes):
    """solve_system! with Time{CrankNicolson} first: x = 2 x* - x_old, x* the
    Krylov solution and x_old the field it started from -- after the
    smoother when one is set."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    n = m.n_cells
    rng = np.random.default_rng(7)
    terms = [(A.TIME_EULER, 1, 1.0), (A.LAPLACIAN_LINEAR, -1, 0.1)]
    prev = rng.standard_normal(n)
    nz, b = om.discretise_scalar(terms, [], prev, dt=0.01)
    x0 = rng.standard_normal(n)
    xs = np.linalg.solve(dense(om, nz), b)
    setup = A.SolverSetupC(A.SOLVER_CG, A.PRECON_DILU, A.PREC_FP64, 1000, 1e-7, 1.0, 1e-300,
                           1e-15, 0.0, 0, 0.0)
    x, _, _ = om.solve_system(setup, nz, b, x0, cn=True)
    assert np.allclose(x, 2 * xs - x0, rtol=0, atol=1e-11 * np.abs(xs).max())
    x_euler, _, _ = om.solve_system(setup, nz, b, x0, cn=False)
    assert np.allclose(x_euler, xs, rtol=0, atol=1e-11 * np.abs(xs).max())
    setup.smoother_loops, setup.smoother_omega = 3, 2 / 3
    x, _, _ = om.solve_system(setup, nz, b, x0, cn=True)
    xsm = om.smooth_jacobi(nz, b, x0, 3, 2 / 3)
    assert np.allclose(x, 2 * xs - xsm, rtol=0, atol=1e-11 * np.abs(xs).max())


def test_crank_nicolson_assembles_as_euler(A, O, meshes):
    """Time{CrankNicolson}'s coefficients are Time{Euler}'s
    (Discretise_1_schemes.jl:44-62)."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    prev = np.linspace(0.0, 1.0, m.n_cells)
    e = om.discretise_scalar([(A.TIME_EULER, 1, 1.0), (A.LAPLACIAN_LINEAR, -1, 0.1)], [], prev,
                             dt=0.02)
    c = om.discretise_scalar([(A.TIME_CRANK_NICOLSON, 1, 1.0), (A.LAPLACIAN_LINEAR, -1, 0.1)],
                             [], prev, dt=0.02)
    assert np.array_equal(e[0], c[0]) and np.array_equal(e[1], c[1])


def test_steady_bfs_reference_check(x, A, O, meshes):
    """test/0_TEST_CASES/2d_incompressible_laminar_BFS.jl:93 with its
    JacobiSmoother on U and p: after 500 SIMPLE iterations the outlet's
    average velocity is half the inlet's within 0.1 U."""
    api = x.api
    m = meshes("bfs10")
    model, config = steady_bfs_case(api, m)
    setup = api.simple_setup(model, config)
    assert setup.U.smoother_loops == 5 and setup.p.smoother_loops == 5
    out = O.OracleMesh(m, A).simple_run(setup, model.momentum.U, model.momentum.p, 500)
    inlet = np.array([0.5, 0.0, 0.0])
    outlet = patch_face_average(m, out["U"], "outlet")
    assert np.linalg.norm(outlet - 0.5 * inlet) <= 0.1 * 0.5, outlet
    assert np.all(np.isfinite(out["U"])) and np.all(np.isfinite(out["p"]))


def test_docs_bfs_verification_profile(x, A, O, meshes):
    """The docs' verification profile (docs/src/examples/figures/01/
    2D_laminar_BFS_XCALibre.csv, Ux, Uy and p at x = 0.501 from an earlier
    release of the reference): the converged oracle run of the same case
    agrees within 2% of the peak velocity in U and 0.01 in p.  A loose
    end-to-end pin: the release that wrote the file is not known."""
    api = x.api
    m = meshes("bfs10")
    model, config = docs_bfs_case(api, m)
    setup = api.simple_setup(model, config)
    out = O.OracleMesh(m, A).simple_run(setup, model.momentum.U, model.momentum.p, 2000)
    assert out["done"] < 2000  # converged to 1e-7
    prof = bfs_verification_profile()
    assert len(prof) == 20
    cc = m.cell_centre.reshape(-1, 3)
    cells = [int(np.argmin(np.hypot(cc[:, 0] - xp, cc[:, 1] - yp) +
                           1e9 * (np.abs(cc[:, 0] - xp) > 0.005) + 1e9 * (np.abs(cc[:, 1] - yp) > 0.005)))
             for xp, yp in prof[:, :2]]
    peak = np.abs(prof[:, 2]).max()
    assert np.abs(out["U"][0, cells] - prof[:, 2]).max() < 0.02 * peak
    assert np.abs(out["U"][1, cells] - prof[:, 3]).max() < 0.02 * peak
    assert np.abs(out["p"][cells] - prof[:, 4]).max() < 0.01

