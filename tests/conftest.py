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
