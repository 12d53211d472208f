"""xcal_precond_apply_noalloc: the preconditioner apply that allocates no
device memory (its sweeps stage through the library's fixed device array).
It must give xcal_precond_apply's results bit for bit, on consecutive calls
whose second mesh is smaller (the array keeps the first call's contents),
and reject meshes above its 2**20-row limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


def _setup(x, A, ctx, mesh, precon, seed):
    H = x.hip
    rng = np.random.default_rng(seed)
    dm = H.DeviceMesh(ctx, mesh)
    nf = mesh.n_faces
    g = H.Field(dm, A.FACE_SCALAR).upload(0.5 + rng.random(nf))
    md = H.Field(dm, A.FACE_SCALAR).upload(rng.standard_normal(nf) * mesh.face_area * 0.2)
    terms = [(A.DIV_UPWIND, 1, md), (A.LAPLACIAN_LINEAR, -1, g)]
    src = H.Field(dm, A.CELL_SCALAR).upload(rng.standard_normal(mesh.n_cells))
    eq = H.Equation(dm, False)
    phi = H.Field(dm, A.CELL_SCALAR)
    eq.discretise(terms, [(src, 1)], phi)
    bcs = A.bc_array([(A.BC_DIRICHLET, i, 1.0) for i in range(len(mesh.patch_names))])
    eq.apply_bcs(terms, bcs, phi)
    eq.precond_update(precon)
    r = H.Field(dm, A.CELL_SCALAR).upload(rng.standard_normal(mesh.n_cells))
    return dm, eq, r


@pytest.mark.parametrize("precon", ["PRECON_DILU", "PRECON_JACOBI"])
def test_noalloc_apply_equals_precond_apply(x, A, ctx, precon):
    """A 110 592-cell random box (one-launch cone sweeps), then a smaller
    8 000-cell box (level launches): each no-allocation apply equals the
    allocating apply on the same equation bit for bit."""
    H = x.hip
    pre = getattr(A, precon)
    for cells, seed in ((48, 3), (20, 4)):
        mesh = x.mesh.box3d(cells, cells, cells, jitter=0.1)
        dm, eq, r = _setup(x, A, ctx, mesh, pre, seed)
        z0 = H.Field(dm, A.CELL_SCALAR)
        z1 = H.Field(dm, A.CELL_SCALAR).upload(np.full(mesh.n_cells, np.nan))
        eq.precond_apply_noalloc(r, z1)
        eq.precond_apply(r, z0)
        a, b = z1.download(), z0.download()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), cells


def test_noalloc_apply_rejects_meshes_above_its_limit(x, A, ctx):
    H = x.hip
    mesh = x.mesh.box3d(102, 102, 102, jitter=0.0, seed_perm=0)  # 1 061 208 rows > 2**20
    dm = H.DeviceMesh(ctx, mesh)
    eq = H.Equation(dm, False)
    r = H.Field(dm, A.CELL_SCALAR)
    z = H.Field(dm, A.CELL_SCALAR)
    with pytest.raises(H.XcalError, match="at most 1048576"):
        eq.precond_apply_noalloc(r, z)
