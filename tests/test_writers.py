"""VTK result writers (src/IOFormats/VTK/VTK_writer.jl, VTK_writer_3D.jl), CPU only.

What is pinned, and how:
* every number reads back to the exact double written, and its digits are
  the shortest that do (the same digits as Python's repr, an independent
  shortest-round-trip implementation);
* the 2D file's lines, counts and cell types against the mesh arrays;
* the 3D polyhedra by geometry: every face of every cell (boundary faces
  included) is listed with its own nodes, ordered so that its normal points
  out of the cell, the face area vectors of every cell sum to zero, and the
  offsets and faceoffsets are the running sums of what precedes them.
Julia's switch between plain and exponent form (1e-4 <= |x| < 1e6 plain) is
stated in include/xcal_io.h; no reference fixture holds writer output, so
that text form is parity unpinned.
"""
import xml.etree.ElementTree as ET

import numpy as np
import pytest


@pytest.fixture(scope="module")
def W(x):
    return x.writers


def repr_digits(v):
    """(sign, significant digits) of Python's shortest repr of v."""
    s = repr(float(v))
    mant = s.lstrip("-").partition("e")[0]
    ip, _, fp = mant.partition(".")
    return s.startswith("-"), (ip + fp).lstrip("0").rstrip("0")


def test_format_double_is_shortest_round_trip(W):
    rng = np.random.default_rng(7)
    vals = [float(v) for v in rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000)]
    vals += [0.1, 1 / 3, 2.0 ** 53, 5e-324, 1.7976931348623157e308, 2.2250738585072014e-308,
             1e-4, 9.999e-5, 1e6, 999999.0, 1234567.0, 123456.7, 1e23, 4.35, 0.3]
    for v in vals:
        t = W.format_double(v)
        assert float(t) == v, (v, t)
        sgn, dig = repr_digits(v)
        m = t.lstrip("-")
        mant, _, e = m.partition("e")
        e = int(e) if e else 0
        ip, _, fp = mant.partition(".")
        d = (ip + fp).lstrip("0").rstrip("0")
        assert d == dig and t.startswith("-") == sgn, (v, t)
        plain = 1e-4 <= abs(v) < 1e6
        assert ("e" not in t) == plain, (v, t)
        if "e" in t:
            assert len(ip) == 1 and fp and "+" not in t, t
        else:
            assert fp, t  # always a fractional part, "100.0"


def test_format_double_special_values(W):
    assert [W.format_double(v) for v in (0.0, -0.0, np.nan, np.inf, -np.inf)] == \
        ["0.0", "-0.0", "NaN", "Inf", "-Inf"]
    assert [W.format_double(v) for v in (1e-5, 1e6, 1.5e-7, 100.0, 0.0001)] == \
        ["1.0e-5", "1.0e6", "1.5e-7", "100.0", "0.0001"]


def nums(lines):
    return np.array([[float(t) for t in ln.split()] for ln in lines])


@pytest.mark.parametrize("name,ctype", [("quad40", 9), ("trig40", 5), ("laplace3", 9)])
def test_vtk2d_file(W, meshes, tmp_path, name, ctype):
    m = meshes(name)
    n = m.n_cells
    rng = np.random.default_rng(1)
    U = rng.standard_normal((3, n))
    p = rng.standard_normal(n) * 1e-7
    T = rng.standard_normal((3, 3, n))
    w = W.initialise_writer(W.VTK(), m)
    path = W.write_results(40, 40, m, w, None, ("U", U), ("p", p), ("tau", T),
                           directory=tmp_path)
    assert path.endswith("iteration_40.vtk")
    L = (tmp_path / "iteration_40.vtk").read_text().split("\n")
    assert L[:4] == ["# vtk DataFile Version 3.0", "XCALibre.jl simulation data", "ASCII",
                     "DATASET UNSTRUCTURED_GRID"]
    assert L[4] == f"POINTS {m.n_nodes} double"
    i = 5
    assert np.array_equal(nums(L[i:i + m.n_nodes]), m.node_coords.reshape(-1, 3))
    i += m.n_nodes
    cnr = m.cell_nodes_range.reshape(-1, 2)
    sizes = cnr[:, 1] - cnr[:, 0] + 1
    assert L[i] == f"CELLS {n} {sizes.sum() + n}"
    for c in range(n):
        ids = m.cell_nodes[cnr[c, 0] - 1:cnr[c, 1]] - 1
        assert L[i + 1 + c] == f"{sizes[c]}  " + " ".join(map(str, ids))
    i += 1 + n
    assert L[i] == f"CELL_TYPES {n}" and set(L[i + 1:i + 1 + n]) == {str(ctype)}
    i += 1 + n
    assert L[i:i + 2] == [f"CELL_DATA {n}", "VECTORS U double"]
    assert np.array_equal(nums(L[i + 2:i + 2 + n]), U.T)
    i += 2 + n
    assert L[i:i + 2] == ["SCALARS p double 1", "LOOKUP_TABLE CellColors"]
    assert np.array_equal(nums(L[i + 2:i + 2 + n])[:, 0], p)
    i += 2 + n
    assert L[i] == "TENSORS tau double"
    got = nums(L[i + 1:i + 1 + 3 * n]).reshape(n, 3, 3)
    assert np.array_equal(got, T.transpose(2, 0, 1))
    assert L[i + 1 + 3 * n:] == [""]


def newell(P):
    """Area vector of a polygon (Newell's formula): half the sum of cross
    products of consecutive vertices."""
    return 0.5 * np.cross(P, np.roll(P, -1, axis=0)).sum(axis=0)


@pytest.mark.parametrize("name", ["cavity_hex", "box12"])
def test_vtu3d_file_polyhedra(W, meshes, tmp_path, name):
    m = meshes(name)
    n = m.n_cells
    X = m.node_coords.reshape(-1, 3)
    rng = np.random.default_rng(2)
    U = rng.standard_normal((3, n))
    k = rng.random(n)
    w = W.initialise_writer(W.VTK(), m)
    W.write_results(7, 0.25, m, w, None, ("U", U), ("k", k), suffix="_post", directory=tmp_path)
    text = (tmp_path / "time_7_post.vtu").read_text()
    assert text.startswith('<?xml version="1.0"?>\n <VTKFile type="UnstructuredGrid" '
                           'version="1.0" byte_order="LittleEndian">\n  <UnstructuredGrid>\n'
                           f'   <Piece NumberOfPoints="{m.n_nodes}" NumberOfCells="{n}">\n')
    assert text.endswith("    </CellData>\n   </Piece>\n  </UnstructuredGrid>\n </VTKFile>\n")
    piece = ET.fromstring(text).find("UnstructuredGrid/Piece")
    arrays = {a.get("Name"): a.text.split() for a in piece.iter("DataArray")}
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=float).reshape(-1, 3)
    assert np.array_equal(pts, X)
    cnr = m.cell_nodes_range.reshape(-1, 2)
    conn = np.array(arrays["connectivity"], dtype=np.int64)
    offs = np.array(arrays["offsets"], dtype=np.int64)
    assert np.array_equal(conn, m.cell_nodes - 1)
    assert np.array_equal(offs, np.cumsum(cnr[:, 1] - cnr[:, 0] + 1))
    assert arrays["types"] == ["42"] * n
    # faces: per cell, the count then (nnodes, ids...) for each face
    owners = m.face_owner.reshape(-1, 2)
    fnr = m.face_nodes_range.reshape(-1, 2)
    fc = m.face_centre.reshape(-1, 3)
    cc = m.cell_centre.reshape(-1, 3)
    cell_faces = [[] for _ in range(n)]
    for f, (a, b) in enumerate(owners):
        cell_faces[a - 1].append(f)
        if b != a:
            cell_faces[b - 1].append(f)
    s = iter(int(t) for t in arrays["faces"])
    fo_expect, total = [], 0
    for c in range(n):
        nf = next(s)
        assert nf == len(cell_faces[c])
        area = np.zeros(3)
        size = 1 + nf
        for f in cell_faces[c]:
            nn = next(s)
            ids = np.array([next(s) for _ in range(nn)])
            mesh_ids = m.face_nodes[fnr[f, 0] - 1:fnr[f, 1]] - 1
            assert list(ids) in (list(mesh_ids), list(mesh_ids[::-1])), (c, f)
            a = newell(X[ids])
            assert np.dot(a, fc[f] - cc[c]) > 0, (c, f)  # outward
            area += a
            size += nn
        scale = max(np.abs(X).max(), 1e-300) ** 2
        assert np.abs(area).max() <= 1e-12 * scale, (c, area)  # closed polyhedron
        total += size
        fo_expect.append(total)
    assert next(s, None) is None
    assert np.array_equal(np.array(arrays["faceoffsets"], dtype=np.int64), fo_expect)
    assert np.array_equal(np.array(arrays["U"], dtype=float).reshape(n, 3), U.T)
    assert np.array_equal(np.array(arrays["k"], dtype=float), k)
    cd = piece.find("CellData")
    assert [(a.get("type"), a.get("NumberOfComponents")) for a in cd] == [("Float32", "3"),
                                                                         ("Float32", None)]


def test_result_names_and_errors(W, meshes, tmp_path):
    m2, m3 = meshes("quad40"), meshes("cavity_hex")
    w2, w3 = W.initialise_writer(W.VTK(), m2), W.initialise_writer(W.VTK(), m3)
    assert w2.name(12, 12) == "iteration_12.vtk"
    assert w2.name(12, 12.0, "_postprocessed") == "iteration_12_postprocessed.vtk"
    assert w3.name(5, 0.005) == "time_5.vtu"
    with pytest.raises(TypeError, match="ScalarField, VectorField or TensorField"):
        W.write_results(1, 1, m2, w2, None, ("U", np.zeros((2, m2.n_cells))))
    with pytest.raises(RuntimeError, match="cannot open"):
        W.write_results(1, 1, m2, w2, None, ("p", np.zeros(m2.n_cells)),
                        directory=tmp_path / "missing")
    with pytest.raises(ValueError, match="writer and mesh differ"):
        W.write_results(1, 1, m3, w2, None)


def foam_list(path):
    """The entries between the outer ( and ) of an OpenFOAM list file."""
    L = path.read_text().split("\n")
    i = L.index("(")
    n = int(L[i - 1])
    j = len(L) - 1 - L[::-1].index(")")
    assert j - i - 1 == n
    return L[i + 1:j]


@pytest.mark.parametrize("name", ["cavity_hex", "box12"])
def test_foam_case_reads_back(W, x, meshes, tmp_path, name):
    """initialise_writer(OpenFOAM(), mesh) (OpenFOAM_writer.jl:1-185): the
    written polyMesh, read back by FOAM3D_mesh, gives the same mesh --
    topology exactly, points to the 6 digits of the reference's %g."""
    m = meshes(name)
    W.initialise_writer(W.OpenFOAM(), m, directory=tmp_path / "case")
    assert (tmp_path / "case" / "XCALibre.foam").read_bytes() == b""
    pm = tmp_path / "case" / "constant" / "polyMesh"
    faces = foam_list(pm / "faces")
    assert faces[0].endswith(" )") and faces[0].split("(")[0] == "4"
    owner = foam_list(pm / "owner")
    assert len(owner) == m.n_faces and len(foam_list(pm / "neighbour")) == m.n_ifaces
    assert f'note        "nPoints: {m.n_nodes} nCells: {m.n_cells} nFaces: {m.n_faces} ' \
           f'nInternalFaces: {m.n_ifaces}";' in (pm / "owner").read_text()
    r = x.mesh.FOAM3D_mesh(pm)
    assert (r.n_cells, r.n_faces, r.n_bfaces, r.n_nodes) == (m.n_cells, m.n_faces, m.n_bfaces,
                                                             m.n_nodes)
    assert r.patch_names == m.patch_names
    for k in ("face_nodes_range", "face_nodes", "face_owner", "patch_range", "cell_faces",
              "cell_nodes", "cell_nodes_range", "cell_neighbours", "boundary_cellsID"):
        assert np.array_equal(getattr(r, k), getattr(m, k)), k
    assert np.allclose(r.node_coords, m.node_coords, rtol=1e-5, atol=0)
    txt = [ln.strip("()").split() for ln in foam_list(pm / "points")]
    assert [[f"{v:g}" for v in row] for row in m.node_coords.reshape(-1, 3)] == txt


def test_foam_fields_and_boundary_entries(W, x, meshes, tmp_path):
    """write_results for the OpenFOAM writer (OpenFOAM_writer.jl:191-273) and
    the boundary entries of _foam_boundary_entry (:277-359)."""
    api = x.api
    m = meshes("box12")
    names = m.patch_names
    assert len(names) == 3
    BCs = api.assign(region=m, U=[api.Wall(names[0], [1, 0, 0]), api.Dirichlet(names[1], [0, 2.5, 0]),
                                  api.Empty(names[2])],
                     p=[api.Dirichlet(names[0], 1e-7), api.Wall(names[1]), api.Symmetry(names[2])],
                     k=[api.KWallFunction(names[0]), api.Neumann(names[1], 0.0),
                        api.Extrapolated(names[2])])
    w = W.initialise_writer(W.OpenFOAM(), m, directory=tmp_path)
    rng = np.random.default_rng(3)
    U, p, k = rng.standard_normal((3, m.n_cells)), rng.standard_normal(m.n_cells), rng.random(m.n_cells)
    assert W.write_results(20, 20, m, w, BCs, ("U", U), ("p", p), ("k", k)) == f"{tmp_path}/20"
    assert W.write_results(4, 0.125, m, w, BCs, ("p", p)) == f"{tmp_path}/0.12500000"
    d = tmp_path / "20"
    Ut = (d / "U").read_text()
    assert Ut.startswith("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
                         "    class       volVectorField;\n    location    \"20\";\n"
                         "    object      U;\n}\n\ninternalField   nonuniform List<vector>\n")
    body = Ut.split("(\n", 1)[1].split("\n);\n")[0].split("\n")
    assert np.array_equal(np.array([[float(t) for t in ln.strip("()").split()] for ln in body]), U.T)
    assert Ut.endswith(
        "boundaryField\n{\n"
        f"\t{names[0]}\n\t{{\n\t\ttype fixedValue;\n\t\tvalue uniform (1.0 0.0 0.0);\n\t}}\n\n"
        f"\t{names[1]}\n\t{{\n\t\ttype fixedValue;\n\t\tvalue uniform (0.0 2.5 0.0);\n\t}}\n\n"
        f"\t{names[2]}\n\t{{\n\t\ttype empty;\n\t}}\n\n}}\n")
    pt = (d / "p").read_text()
    assert "class       volScalarField;" in pt and "internalField   nonuniform List<scalar>\n" in pt
    body = pt.split("(\n", 1)[1].split("\n);\n")[0].split("\n")
    assert np.array_equal(np.array([float(t) for t in body]), p)
    zg = "\t{\n\t\ttype zeroGradient;\n\t}\n\n"
    assert pt.endswith("boundaryField\n{\n"
                       f"\t{names[0]}\n\t{{\n\t\ttype fixedValue;\n\t\tvalue uniform 1.0e-7;\n\t}}\n\n"
                       f"\t{names[1]}\n{zg}\t{names[2]}\n{zg}}}\n")
    assert (d / "k").read_text().endswith(
        "boundaryField\n{\n" + "".join(f"\t{n}\n{zg}" for n in names) + "}\n")
    assert 'location    "4";' in (tmp_path / "0.12500000" / "p").read_text()
    with pytest.raises(TypeError, match="ScalarField or VectorField"):
        W.write_results(1, 1, m, w, BCs, ("U", np.zeros((9, m.n_cells))))
    with pytest.raises(RuntimeError, match="only be used for 3D"):
        W.initialise_writer(W.OpenFOAM(), meshes("quad40"), directory=tmp_path / "2d")
