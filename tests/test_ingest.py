"""Matrix connectivity of a mesh (CSR pattern, slot maps, DILU levels), the
host form: against the oracle's CSR and against what the definitions fix.

The device form must equal this one element for element (test_gpu_ingest.py).
Definitions checked here, independently of the builder's own loops:
  * every row holds its cell and its neighbours, ascending and unique
    (the oracle's CSR, from sparse_matrix_connectivity + sparsecsr);
  * colval[slot] is the column each slot map names, inside the right row;
  * forward level = 0 without lower neighbours, else 1 + the maximum level of
    the lower neighbours (the longest chain of dependencies); backward the
    same over upper neighbours."""
import numpy as np
import pytest

MESHES = ["laplace3", "cavity40", "box12", "cavity_hex", "trig40", "chain5"]


def _rows(rp):
    return np.repeat(np.arange(len(rp) - 1), np.diff(rp))


def check_connectivity(K, mesh, orp=None, ocv=None):
    rp, cv = K["rowptr"], K["colval"]
    n = mesh.n_cells
    rows = _rows(rp)
    if orp is not None:
        assert np.array_equal(rp, orp) and np.array_equal(cv, ocv)
    # ascending unique columns per row, diagonal present
    same_row = rows[1:] == rows[:-1]
    assert np.all(cv[1:][same_row] > cv[:-1][same_row])
    d = K["diag"]
    assert np.array_equal(cv[d], np.arange(n)) and np.array_equal(rows[d], np.arange(n))
    # cell-face pairs in faces_range order
    rng = np.asarray(mesh.cell_faces_range).reshape(-1, 2)
    nb = np.asarray(mesh.cell_neighbours) - 1
    cell_of = np.concatenate([np.full(b - a + 1, c) for c, (a, b) in enumerate(rng)])
    k_of = np.concatenate([np.arange(a - 1, b) for a, b in rng])
    s = K["cf_slot"]
    assert np.array_equal(cv[s], nb[k_of]) and np.array_equal(rows[s], cell_of)
    # transpose slots
    t = K["trans"]
    assert np.all(t >= 0)  # the pattern is symmetric
    assert np.array_equal(rows[t], cv) and np.array_equal(cv[t], rows)
    # owner-row slot of every internal face
    fo = np.asarray(mesh.face_owner).reshape(-1, 2) - 1
    nbf = mesh.n_bfaces
    fs = K["face_slot"]
    assert np.all(fs[:nbf] == -1)
    assert np.array_equal(rows[fs[nbf:]], fo[nbf:, 0]) and np.array_equal(cv[fs[nbf:]], fo[nbf:, 1])
    # levels: the longest-chain recursion, over lower and over upper entries
    for lev, part in ((K["flev"], cv < rows), (K["blev"], cv > rows)):
        dep = np.full(n, -1)
        np.maximum.at(dep, rows[part], lev[cv[part]])
        assert np.array_equal(lev, dep + 1)
    assert K.levels == (int(K["flev"].max()) + 1, int(K["blev"].max()) + 1)


@pytest.mark.parametrize("name", MESHES)
def test_host_connectivity_matches_oracle_and_definitions(x, A, O, meshes, name):
    mesh = meshes(name)
    K = x.hip.Connectivity(None, mesh, on_device=False)
    orp, ocv = O.OracleMesh(mesh, A).csr()
    check_connectivity(K, mesh, orp, ocv)


def test_definitions_catch_a_wrong_level(x, A, meshes):
    """The level check fails on a plausible mistake (levels of a row taken as
    its lower neighbour count instead of the longest chain)."""
    mesh = meshes("cavity40")
    K = x.hip.Connectivity(None, mesh, on_device=False)
    rows = _rows(K["rowptr"])
    wrong = np.bincount(rows[K["colval"] < rows], minlength=mesh.n_cells).astype(np.int32)
    K.arrays["flev"] = wrong
    with pytest.raises(AssertionError):
        check_connectivity(K, mesh)
