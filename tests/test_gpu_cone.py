"""Cone form of the DILU sweeps.

On a randomly numbered mesh of at least 65 536 rows each sweep is one launch:
the workgroup of a block of cells sweeps the block's rows and every row they
depend on (the block's dependency cone) in LDS, computing the other blocks'
rows again instead of waiting for them.  Every row uses the reference's
formulas in its column order (forward_substitution!, backward_substitution!,
preconditioners_1_DILU.jl:84-114), so the preconditioner applies bit for bit
as the oracle and the level launches do; Krylov solves and SIMPLE runs differ
from the level-launch form only by the order of their dot-product sums.
"""
import numpy as np
import pytest

from conftest import channel_case, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


@pytest.fixture(scope="module")
def box48(x):
    return x.mesh.box3d(48, 48, 48, jitter=0.1)  # 110 592 cells, random numbering


@pytest.fixture(scope="module")
def box48n(x):
    # natural (lexicographic) numbering: a dependency chain crosses the box
    # (142 levels), so the cones of all levels do not fit and the sweeps run
    # in bands of levels
    return x.mesh.box3d(48, 48, 48, jitter=0.1, seed_perm=0)


def _mesh(x, ctx, mesh, monkeypatch, **env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    dm = x.hip.DeviceMesh(ctx, mesh)
    for k in env:
        monkeypatch.delenv(k)
    return dm


def _equation(x, A, dm, mesh, rng, sym):
    H = x.hip
    nf = mesh.n_faces
    g = H.Field(dm, A.FACE_SCALAR).upload(0.5 + rng.random(nf))
    mdh = np.zeros(nf) if sym else rng.standard_normal(nf) * mesh.face_area * 0.2
    md = H.Field(dm, A.FACE_SCALAR).upload(mdh)
    terms = [(A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)]
    src = H.Field(dm, A.CELL_SCALAR).upload(rng.standard_normal(mesh.n_cells))
    eq = H.Equation(dm, False)
    phi = H.Field(dm, A.CELL_SCALAR)
    eq.discretise(terms, [(src, 1)], phi)
    bcs = A.bc_array([(A.BC_DIRICHLET, i, 1.0) for i in range(len(mesh.patch_names))])
    eq.apply_bcs(terms, bcs, phi)
    return eq


def test_cone_form_is_chosen_on_random_numbering(x, ctx, box48, box48n, monkeypatch):
    dm = _mesh(x, ctx, box48, monkeypatch)
    rows, nblk, fr, br = dm.cone_form()
    n = box48.n_cells
    assert rows == 2048 and nblk == (n + rows - 1) // rows
    assert n < fr < 3 * n and n < br < 3 * n
    assert dm.cone_bands() == (1, 1)  # every level in one launch
    no = _mesh(x, ctx, box48, monkeypatch, XCAL_NO_CONE="1")
    assert no.cone_form() == (0, 0, 0, 0) and no.cone_bands() == (0, 0)
    # natural numbering: bands of at least 4 levels, far fewer launches than levels
    dn = _mesh(x, ctx, box48n, monkeypatch)
    fb, bb = dn.cone_bands()
    lf, lb = dn.levels
    print(f"natural 48^3: {lf} levels, {fb} forward and {bb} backward bands")
    assert dn.cone_form()[0] > 0 and 1 < fb <= lf // 4 and 1 < bb <= lb // 4


@pytest.mark.parametrize("rows,deep,numbering", [
    ("2048", "0", "random"), ("1024", "1", "random"), ("256", "0", "random"),
    ("2048", "0", "natural"), ("512", "1", "natural")])
def test_cone_preconditioner_bit_exact(x, A, O, ctx, box48, box48n, monkeypatch, rows, deep,
                                       numbering):
    """z = M^-1 r through both sweeps equals the level launches and the
    oracle's sequential loops bit for bit, for blocks of 2048, 1024, 512 and
    256 rows, both prefetch depths, and both numberings (natural numbering
    sweeps in bands of levels, each block reading the rows of earlier bands
    from global memory)."""
    H = x.hip
    rng = np.random.default_rng(11)
    box48 = box48 if numbering == "random" else box48n
    dm = _mesh(x, ctx, box48, monkeypatch, XCAL_CONE_ROWS=rows, XCAL_CONE_DEEP=deep)
    ref = _mesh(x, ctx, box48, monkeypatch, XCAL_NO_CONE="1")
    assert dm.cone_form()[0] == int(rows)
    assert (dm.cone_bands() == (1, 1)) == (numbering == "random")
    r_h = rng.standard_normal(box48.n_cells)
    out = []
    for m in (dm, ref):
        eq = _equation(x, A, m, box48, np.random.default_rng(5), sym=False)
        eq.precond_update(A.PRECON_DILU)
        z = H.Field(m, A.CELL_SCALAR)
        eq.precond_apply(H.Field(m, A.CELL_SCALAR).upload(r_h), z)
        out.append(z.download())
    assert np.array_equal(out[0], out[1])
    om = O.OracleMesh(box48, A)
    rng5 = np.random.default_rng(5)
    nf = box48.n_faces
    gam = 0.5 + rng5.random(nf)
    mdh = rng5.standard_normal(nf) * box48.face_area * 0.2
    src = rng5.standard_normal(box48.n_cells)
    oterms = [(A.DIV_UPWIND, 1, mdh), (A.LAPLACIAN_LINEAR, -1, gam)]
    onz, ob = om.discretise_scalar(oterms, [(src, 1)])
    bcs = A.bc_array([(A.BC_DIRICHLET, i, 1.0) for i in range(len(box48.patch_names))])
    om.apply_bcs(oterms, bcs, -1, np.zeros(box48.n_cells), onz, ob)
    oz = om.dilu_apply(onz, om.dilu_diagonal(onz), r_h)
    assert np.array_equal(out[0], oz)


@pytest.mark.parametrize("numbering", ["random", "natural"])
@pytest.mark.parametrize("solver", ["SOLVER_CG", "SOLVER_BICGSTAB"])
@pytest.mark.parametrize("prec", ["PREC_FP64", "PREC_FP32_IR", "PREC_FP32_PRECON"])
def test_cone_krylov_equals_level_form(x, A, ctx, box48, box48n, monkeypatch, solver, prec,
                                       numbering):
    """Solutions to rtol 1e-12 agree to 1e-9 relative (the dot products are
    summed per block instead of per level launch); FP64 iteration counts
    within 5% (CG) and 25% (BiCGSTAB).  Under PREC_FP32_PRECON the cone
    sweeps of CG compute in FP32 and the level form keeps FP64, so the two
    differ by the preconditioner's rounding as well: CG's FP64 residual
    recurrence still reaches the same solution.  BiCGSTAB keeps FP64 sweeps
    under that mode.  Natural numbering runs the cone sweeps in bands."""
    H = x.hip
    box48 = box48 if numbering == "random" else box48n
    sym = solver == "SOLVER_CG"
    res = []
    for env in ({}, {"XCAL_NO_CONE": "1"}):
        m = _mesh(x, ctx, box48, monkeypatch, **env)
        eq = _equation(x, A, m, box48, np.random.default_rng(9), sym)
        eq.precond_update(A.PRECON_DILU)
        phi = H.Field(m, A.CELL_SCALAR)
        setup = A.SolverSetupC(getattr(A, solver), A.PRECON_DILU, getattr(A, prec), 4000, 1e-7,
                               1.0, 1e-30, 1e-12, 0.0)
        it, r, rc = eq.solve(setup, phi)
        assert rc == 0
        res.append((phi.download(), it))
    (a, ia), (b, ib) = res
    assert rel(a, b) < 1e-9
    if prec == "PREC_FP64":
        slack = ib // 20 if sym else ib // 4
        assert abs(ia - ib) <= max(2, slack), (ia, ib)


@pytest.mark.parametrize("prec", ["fp64", "fp32-ir", "fp32-precon"])
def test_cone_simple_channel_equals_level_form(x, ctx, box48, monkeypatch, prec):
    """Five SIMPLE iterations of the channel case (p: CG+DILU to 1e-10, U:
    BiCGSTAB+DILU) with cone and level sweeps agree to 1e-9 in U and p
    (under fp32-precon the level form's sweeps stay FP64)."""
    api = x.api
    out = []
    for env in ({}, {"XCAL_NO_CONE": "1"}):
        m = _mesh(x, ctx, box48, monkeypatch, **env)
        model, config = channel_case(api, box48, iterations=5, precision_p=prec)
        sim = x.hip.Simple(m, api.simple_setup(model, config))
        sim.set_U(model.momentum.U)
        sim.set_p(model.momentum.p)
        sim.run(5)
        out.append((sim.get("U"), sim.get("p")))
    assert rel(out[0][0], out[1][0]) < 1e-9
    assert rel(out[0][1], out[1][1]) < 1e-9


def test_fp32_preconditioner_iterations_and_batch_rejection(x, A, ctx, box48, monkeypatch):
    """The FP32 sweeps inside FP64 CG keep the Krylov space: the iteration
    count to rtol 1e-10 stays within 10% of the FP64 preconditioner's, and
    the solutions agree to 1e-8.  A batch mesh rejects the mode."""
    H = x.hip
    m = _mesh(x, ctx, box48, monkeypatch)
    res = {}
    for prec in ("PREC_FP64", "PREC_FP32_PRECON"):
        eq = _equation(x, A, m, box48, np.random.default_rng(9), sym=True)
        eq.precond_update(A.PRECON_DILU)
        phi = H.Field(m, A.CELL_SCALAR)
        setup = A.SolverSetupC(A.SOLVER_CG, A.PRECON_DILU, getattr(A, prec), 4000, 1e-7,
                               1.0, 1e-30, 1e-10, 0.0)
        it, r, rc = eq.solve(setup, phi)
        assert rc == 0
        res[prec] = (phi.download(), it)
    (a, ia), (b, ib) = res["PREC_FP64"], res["PREC_FP32_PRECON"]
    print(f"CG+DILU to 1e-10: FP64 sweeps {ia} iterations, FP32 sweeps {ib}")
    assert rel(b, a) < 1e-8
    assert abs(ib - ia) <= max(2, ia // 10), (ia, ib)
    cav = [x.mesh.cavity2d(8), x.mesh.cavity2d(12)]
    bm = H.DeviceMesh.batch(ctx, cav)
    eq = H.Equation(bm, False)
    g = H.Field(bm, A.FACE_SCALAR)
    g.fill(1.0)
    phi = H.Field(bm, A.CELL_SCALAR)
    src = H.Field(bm, A.CELL_SCALAR)
    src.fill(1.0)
    terms = [(A.LAPLACIAN_LINEAR, -1, g)]
    eq.discretise(terms, [(src, 1)], phi)
    eq.precond_update(A.PRECON_DILU)
    setup = A.SolverSetupC(A.SOLVER_CG, A.PRECON_DILU, A.PREC_FP32_PRECON, 100, 1e-7, 1.0,
                           1e-30, 1e-8, 0.0)
    with pytest.raises(H.XcalError, match="single meshes"):
        eq.solve(setup, phi)
