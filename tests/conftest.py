"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and run on an MI355X;
everything else runs on the CPU (oracle, mesh builders, host logic, ABI)."""
import gzip
import shutil
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def x():
    return ge.load_package()


@pytest.fixture(scope="session")
def A(x):
    return x._abi


@pytest.fixture(scope="session")
def O():
    return ge.load_oracle()


@pytest.fixture(scope="session")
def fixture_dir(tmp_path_factory):
    """Unpacked copies of the gzipped reference mesh fixtures."""
    d = tmp_path_factory.mktemp("meshes")
    src = GOLDEN / "meshes"
    for gz in src.rglob("*.gz"):
        rel = gz.relative_to(src).with_suffix("")
        out = d / rel
        out.parent.mkdir(parents=True, exist_ok=True)
        with gzip.open(gz, "rb") as f, open(out, "wb") as g:
            shutil.copyfileobj(f, g)
    return d


@pytest.fixture(scope="session")
def meshes(x, fixture_dir):
    """Lazily built, cached meshes by name."""
    cache = {}

    def get(name):
        if name not in cache:
            M = x.mesh
            if name == "laplace3":
                cache[name] = M.UNV2D_mesh(fixture_dir / "laplace_unit_3by3.unv")
            elif name == "quad40":
                cache[name] = M.UNV2D_mesh(fixture_dir / "quad40.unv", 0.001)
            elif name == "trig40":
                cache[name] = M.UNV2D_mesh(fixture_dir / "trig40.unv", 0.001)
            elif name == "cavity_hex":
                cache[name] = M.FOAM3D_mesh(fixture_dir / "OF_cavity_hex", 0.001)
            elif name == "bfs10":
                cache[name] = M.UNV2D_mesh(fixture_dir / "backwardFacingStep_10mm.unv", 0.001)
            elif name == "cavity40":
                cache[name] = M.cavity2d(40)
            elif name == "cavity16":
                cache[name] = M.cavity2d(16)
            elif name == "box12":
                cache[name] = M.box3d(12, 10, 8, jitter=0.1)
            elif name == "box20":
                cache[name] = M.box3d(20, 20, 20, jitter=0.1)
            elif name == "box70":  # 343 000 cells: above the CG split threshold
                cache[name] = M.box3d(70, 70, 70, jitter=0.1)
            elif name == "box12_natural":
                cache[name] = M.box3d(12, 10, 8, jitter=0.0, seed_perm=0)
            elif name == "chain5":  # five quads in a row: a tridiagonal CSR pattern
                cache[name] = M.block_mesh_2d(
                    [(0.0, 0.0, 0.0), (5.0, 0.0, 0.0), (0.0, 1.0, 0.0), (5.0, 1.0, 0.0)],
                    [(1, 2, 5), (3, 4, 5), (1, 3, 1), (2, 4, 1)], [(1, 2, 3, 4)],
                    [("bottom", [1]), ("top", [2]), ("left", [3]), ("right", [4])])
            elif name == "step":
                cache[name] = M.backward_facing_step_2d(0.4, nx_in=12, nx_out=30, ny_low=6, ny_up=8)
            else:
                raise KeyError(name)
        return cache[name]

    return get


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def cavity_case(api, mesh, nu=0.01, iterations=200, precon="DILU", rtol_p=0.1,
                precision_p="fp64"):
    P = getattr(api, precon)
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=nu))
    BCs = api.assign(region=mesh,
                     U=[api.Wall("bottom", [0, 0, 0]), api.Wall("top", [1, 0, 0]),
                        api.Wall("left", [0, 0, 0]), api.Wall("right", [0, 0, 0])],
                     p=[api.Wall("bottom"), api.Wall("top"), api.Wall("left"), api.Wall("right")])
    solvers = dict(U=api.SolverSetup(solver=api.Bicgstab(), preconditioner=P(),
                                     convergence=1e-7, relax=0.7),
                   p=api.SolverSetup(solver=api.Cg(), preconditioner=P(),
                                     convergence=1e-7, relax=0.3, rtol=rtol_p,
                                     precision=precision_p))
    config = api.Configuration(solvers=solvers,
                               schemes=dict(U=api.Schemes(divergence=api.Upwind), p=api.Schemes()),
                               runtime=api.Runtime(iterations=iterations, write_interval=-1),
                               hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config


def transient_bfs_case(api, mesh, steps, dt=0.005, precon="Jacobi", rtol=1e-3):
    """The reference's test/0_TEST_CASES/2d_incompressible_transient_laminar_BFS.jl
    (U 0.5, nu 1e-3, Euler, dt 0.005, Jacobi, relax 1, rtol 1e-3), with CG in
    place of GMRES for the pressure."""
    P = getattr(api, precon)
    u = 0.5
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=1e-3), time=api.Transient())
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [u, 0, 0]), api.Extrapolated("outlet"),
                        api.Wall("wall", [0, 0, 0]), api.Dirichlet("top", [0, 0, 0])],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Extrapolated("wall"), api.Extrapolated("top")])

    def S(s):
        return api.SolverSetup(solver=s, preconditioner=P(), convergence=1e-7, relax=1.0,
                               rtol=rtol)
    config = api.Configuration(
        solvers=dict(U=S(api.Bicgstab()), p=S(api.Cg())),
        schemes=dict(U=api.Schemes(time=api.Euler), p=api.Schemes()),
        runtime=api.Runtime(iterations=steps, write_interval=-1, time_step=dt),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    api.initialise(model.momentum.U, [u, 0, 0])
    return model, config


def patch_face_average(mesh, U, patch):
    """boundary_average(patch, U, BCs.U, config) (Postprocess_functions.jl:157-181)
    for a zero-gradient patch (Extrapolated): face values are the owner
    cells' values, averaged without weights."""
    k = mesh.patch_index(patch)
    f0, f1 = mesh.patch_range[2 * k] - 1, mesh.patch_range[2 * k + 1]
    cells = mesh.boundary_cellsID[f0:f1] - 1
    return np.asarray(U)[:, cells].mean(axis=1)


def channel_case(api, mesh, nu=1e-3, iterations=10, precon="DILU", rtol_p=1e-10,
                 div=None, precision_p="fp64"):
    P = getattr(api, precon)
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=nu))
    api.initialise(model.momentum.U, [1.0, 0.0, 0.0])
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [1, 0, 0]), api.Extrapolated("outlet"),
                        api.Wall("walls", [0, 0, 0])],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Extrapolated("walls")])
    solvers = dict(U=api.SolverSetup(solver=api.Bicgstab(), preconditioner=P(),
                                     convergence=1e-7, relax=0.7),
                   p=api.SolverSetup(solver=api.Cg(), preconditioner=P(),
                                     convergence=1e-7, relax=0.3, rtol=rtol_p,
                                     precision=precision_p))
    config = api.Configuration(
        solvers=solvers,
        schemes=dict(U=api.Schemes(divergence=div or api.Upwind), p=api.Schemes()),
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config


def komega_bfs_case(api, mesh, iterations, precon="Jacobi"):
    """The reference's examples/2D_BFS_highRe.jl settings (k-omega with wall
    functions, U 1.5, nu 1e-3, inlet k 1 and omega 1000) on a step mesh."""
    nu, u, k_in, w_in = 1e-3, 1.5, 1.0, 1000.0
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=nu), turbulence=api.RANS(api.KOmega))
    BCs = api.assign(
        region=mesh,
        U=[api.Dirichlet("inlet", [u, 0, 0]), api.Extrapolated("outlet"),
           api.Wall("wall", [0, 0, 0]), api.Wall("top", [0, 0, 0])],
        p=[api.Neumann("inlet", 0.0), api.Dirichlet("outlet", 0.0), api.Wall("wall"),
           api.Wall("top")],
        k=[api.Dirichlet("inlet", k_in), api.Extrapolated("outlet"), api.KWallFunction("wall"),
           api.KWallFunction("top")],
        omega=[api.Dirichlet("inlet", w_in), api.Extrapolated("outlet"),
               api.OmegaWallFunction("wall"), api.OmegaWallFunction("top")],
        nut=[api.Dirichlet("inlet", k_in / w_in), api.Extrapolated("outlet"),
             api.NutWallFunction("wall"), api.NutWallFunction("top")])
    P = getattr(api, precon)

    def S(s, r, rt):
        return api.SolverSetup(solver=s, preconditioner=P(), convergence=1e-7, relax=r, rtol=rt,
                               atol=1e-10)
    solvers = dict(U=S(api.Bicgstab(), 0.7, 1e-2), p=S(api.Cg(), 0.3, 1e-3),
                   k=S(api.Bicgstab(), 0.7, 1e-2), omega=S(api.Bicgstab(), 0.7, 1e-2))
    schemes = {f: api.Schemes(divergence=api.Upwind) for f in ("U", "p", "k", "omega")}
    config = api.Configuration(
        solvers=solvers, schemes=schemes,
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config


def steady_bfs_case(api, mesh):
    """The reference's test/0_TEST_CASES/2d_incompressible_laminar_BFS.jl:
    U 0.5, nu 1e-3, Linear convection, Symmetry top, JacobiSmoother(loops 5,
    omega 2/3) and Jacobi preconditioners on U (BiCGSTAB, relax 0.8, rtol
    0.1) and p (CG, relax 0.2, rtol 0.01), 500 iterations."""
    u = 0.5
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=1e-3))
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [u, 0, 0]), api.Extrapolated("outlet"),
                        api.Wall("wall", [0, 0, 0]), api.Symmetry("top")],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Extrapolated("wall"), api.Symmetry("top")])
    sm = api.JacobiSmoother(domain=mesh, loops=5, omega=2 / 3)
    solvers = dict(
        U=api.SolverSetup(solver=api.Bicgstab(), smoother=sm, preconditioner=api.Jacobi(),
                          convergence=1e-7, relax=0.8, rtol=1e-1),
        p=api.SolverSetup(solver=api.Cg(), smoother=sm, preconditioner=api.Jacobi(),
                          convergence=1e-7, relax=0.2, rtol=1e-2))
    config = api.Configuration(
        solvers=solvers, schemes=dict(U=api.Schemes(divergence=api.Linear), p=api.Schemes()),
        runtime=api.Runtime(iterations=500, write_interval=500),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    api.initialise(model.momentum.U, [u, 0, 0])
    return model, config


def docs_bfs_case(api, mesh, iterations=2000):
    """The verification case of docs/src/examples/01_2d-isothermal-backward-
    facing-step.md: U 1.5, nu 1e-3, Linear convection, walls top and bottom,
    Jacobi preconditioners, relax 0.7, rtol 1e-4, atol 1e-10."""
    u = 1.5
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=1e-3))
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [u, 0, 0]), api.Extrapolated("outlet"),
                        api.Wall("wall", [0, 0, 0]), api.Wall("top", [0, 0, 0])],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Wall("wall"), api.Wall("top")])

    def S(s):
        return api.SolverSetup(solver=s, preconditioner=api.Jacobi(), convergence=1e-7,
                               relax=0.7, rtol=1e-4, atol=1e-10)
    config = api.Configuration(
        solvers=dict(U=S(api.Bicgstab()), p=S(api.Cg())),
        schemes=dict(U=api.Schemes(divergence=api.Linear), p=api.Schemes()),
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    api.initialise(model.momentum.U, [u, 0, 0])
    return model, config


def bfs_verification_profile():
    """Rows (x, y, Ux, Uy, p) of tests/golden/2D_laminar_BFS_XCALibre.csv,
    the docs' profile at x = 0.501 (ParaView's line probe; its
    vtkValidPointMask column holds NUL bytes, and points outside the domain
    are nan)."""
    import csv
    import io
    text = (GOLDEN / "2D_laminar_BFS_XCALibre.csv").read_bytes().replace(b"\x00", b"").decode()
    rows = list(csv.reader(io.StringIO(text)))[1:]
    out = [(float(r[6]), float(r[7]), float(r[0]), float(r[1]), float(r[3])) for r in rows
           if r[0] != "nan"]
    return np.array(out)

