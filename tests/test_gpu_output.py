"""Result output during device runs.

run(model, config) writes VTK files (or an OpenFOAM case with
output=OpenFOAM()) where the reference's run! does: after
every iteration that is a multiple of write_interval, and at the SIMPLE
iteration that converges (Solvers_1_SIMPLE.jl:234-258,
Solvers_2_PISO.jl:205-208).  The iterations between two writes run as one
device call on state that stays resident, so the pinned properties are:
the chunked run is bit-identical to one uninterrupted run (fields and
residual histories), each file holds bit-exactly the fields of an
uninterrupted run stopped at that iteration, and the file names follow the
reference's iteration_N / time_N rule."""
import dataclasses
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import cavity_case, channel_case, transient_bfs_case

pytestmark = pytest.mark.gpu


def interval(config, wi, iterations=None):
    rt = dataclasses.replace(config.runtime, write_interval=wi,
                             iterations=iterations or config.runtime.iterations)
    return dataclasses.replace(config, runtime=rt)


def vtk_cell_data(path, n):
    L = path.read_text().split("\n")
    out = {}
    for i, ln in enumerate(L):
        if ln.startswith("VECTORS "):
            out[ln.split()[1]] = np.array([[float(t) for t in r.split()] for r in L[i + 1:i + 1 + n]]).T
        elif ln.startswith("SCALARS "):
            out[ln.split()[1]] = np.array([float(r) for r in L[i + 2:i + 2 + n]])
    return out


def vtu_cell_data(path):
    piece = ET.fromstring(path.read_text()).find("UnstructuredGrid/Piece")
    out = {}
    for a in piece.find("CellData"):
        v = np.array(a.text.split(), dtype=float)
        out[a.get("Name")] = v.reshape(-1, 3).T if a.get("NumberOfComponents") == "3" else v
    return out


def run_case(api, make, wi, iterations, directory=None):
    model, config = make()
    r = api.run(model, interval(config, wi, iterations), directory=directory)
    return model, r


def test_simple_writes_at_interval_and_matches_one_run(x, meshes, tmp_path):
    api = x.api
    mesh = meshes("cavity16")

    def make():
        return cavity_case(api, mesh, iterations=10)
    ref, r_ref = run_case(api, make, -1, 10)
    got, r_got = run_case(api, make, 4, 10, tmp_path)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["iteration_4.vtk", "iteration_8.vtk"]
    assert np.array_equal(got.momentum.U, ref.momentum.U)
    assert np.array_equal(got.momentum.p, ref.momentum.p)
    for k in r_ref:
        assert np.array_equal(r_got[k], r_ref[k]), k
    at8, _ = run_case(api, make, -1, 8)
    d = vtk_cell_data(tmp_path / "iteration_8.vtk", mesh.n_cells)
    assert np.array_equal(d["U"], at8.momentum.U) and np.array_equal(d["p"], at8.momentum.p)


def test_simple_writes_at_convergence(x, meshes, tmp_path):
    api = x.api
    mesh = meshes("cavity16")

    def make():
        model, config = cavity_case(api, mesh, iterations=200)
        solvers = {k: dataclasses.replace(v, convergence=2e-3) for k, v in config.solvers.items()}
        return model, dataclasses.replace(config, solvers=solvers)
    got, r = run_case(api, make, 1000, 200, tmp_path)
    n = len(r["p"])
    assert 1 < n < 200  # stopped on convergence, before any interval write
    assert [p.name for p in tmp_path.iterdir()] == [f"iteration_{n}.vtk"]
    d = vtk_cell_data(tmp_path / f"iteration_{n}.vtk", mesh.n_cells)
    assert np.array_equal(d["U"], got.momentum.U) and np.array_equal(d["p"], got.momentum.p)


def test_piso_writes_time_steps_and_matches_one_run(x, meshes, tmp_path):
    api = x.api
    mesh = meshes("bfs10")

    def make():
        return transient_bfs_case(api, mesh, 6)
    ref, r_ref = run_case(api, make, -1, 6)
    got, r_got = run_case(api, make, 3, 6, tmp_path)
    # time = 3 * 0.005 differs from the iteration count: time_N names
    assert sorted(p.name for p in tmp_path.iterdir()) == ["time_3.vtk", "time_6.vtk"]
    assert np.array_equal(got.momentum.U, ref.momentum.U)
    assert np.array_equal(got.momentum.p, ref.momentum.p)
    for k in r_ref:
        assert np.array_equal(r_got[k], r_ref[k]), k
    d = vtk_cell_data(tmp_path / "time_6.vtk", mesh.n_cells)
    assert np.array_equal(d["U"], ref.momentum.U) and np.array_equal(d["p"], ref.momentum.p)


def test_3d_run_writes_vtu(x, meshes, tmp_path):
    api = x.api
    mesh = meshes("box12")

    def make():
        return channel_case(api, mesh, iterations=6)
    ref, _ = run_case(api, make, -1, 6)
    at3, _ = run_case(api, make, -1, 3)
    got, _ = run_case(api, make, 3, 6, tmp_path)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["iteration_3.vtu", "iteration_6.vtu"]
    assert np.array_equal(got.momentum.U, ref.momentum.U)
    for name, m in (("iteration_3.vtu", at3), ("iteration_6.vtu", ref)):
        d = vtu_cell_data(tmp_path / name)
        assert np.array_equal(d["U"], m.momentum.U) and np.array_equal(d["p"], m.momentum.p)


def foam_internal(path):
    body = path.read_text().split("(\n", 1)[1].split("\n);\n")[0].split("\n")
    v = np.array([[float(t) for t in ln.strip("()").split()] for ln in body])
    return v.T if v.shape[1] == 3 else v[:, 0]


def test_3d_run_writes_openfoam_case(x, meshes, tmp_path):
    api = x.api
    mesh = meshes("box12")

    def make():
        return channel_case(api, mesh, iterations=6)
    at3, _ = run_case(api, make, -1, 3)
    model, config = make()
    api.run(model, interval(config, 3, 6), output=api.OpenFOAM(), directory=tmp_path)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["3", "6", "XCALibre.foam", "constant"]
    back = x.mesh.FOAM3D_mesh(tmp_path / "constant" / "polyMesh")
    assert np.array_equal(back.face_nodes, mesh.face_nodes)
    for d, m in (("3", at3), ("6", model)):
        assert np.array_equal(foam_internal(tmp_path / d / "U"), m.momentum.U)
        assert np.array_equal(foam_internal(tmp_path / d / "p"), m.momentum.p)
