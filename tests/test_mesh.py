"""Host mesh builders against the reference's own mesh fixtures and the
invariants every finite-volume mesh must satisfy (CPU only)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

COUNTS = json.load(open(GOLDEN / "mesh_counts.json"))


@pytest.mark.parametrize("name,key", [("quad40", "quad40"), ("trig40", "trig40"),
                                      ("cavity_hex", "OF_cavity_hex")])
def test_reference_mesh_counts(meshes, name, key):
    """test_mesh_conversion.jl:33,41,57: cells/faces/nodes of the fixtures."""
    m = meshes(name)
    assert [m.n_cells, m.n_faces, m.n_nodes] == COUNTS[key]
    kind = "3D" if name == "cavity_hex" else "2D"
    assert str(m) == f"{kind} Mesh with:\n-> {COUNTS[key][0]} cells\n-> {COUNTS[key][1]} faces\n" \
                     f"-> {COUNTS[key][2]} nodes"


ALL = ["laplace3", "quad40", "trig40", "cavity_hex", "cavity40", "box12", "box12_natural", "step"]


@pytest.mark.parametrize("name", ALL)
def test_boundary_first_and_owner_order(meshes, name):
    m = meshes(name)
    own = m.face_owner.reshape(-1, 2)
    nbf = m.n_bfaces
    assert np.all(own[:nbf, 0] == own[:nbf, 1])             # boundary: (c, c)
    assert np.all(own[nbf:, 0] < own[nbf:, 1])              # internal: owner < neighbour
    assert np.array_equal(m.boundary_cellsID, own[:nbf, 0])
    pr = m.patch_range.reshape(-1, 2)
    assert pr[0, 0] == 1 and pr[-1, 1] == nbf and np.all(pr[1:, 0] == pr[:-1, 1] + 1)
    # every internal face appears once from each side with opposite nsign
    cf, cn, cs = m.cell_faces, m.cell_neighbours, m.cell_nsign
    assert len(cf) == 2 * m.n_ifaces
    rng = m.cell_faces_range.reshape(-1, 2)
    cell_of = np.repeat(np.arange(1, m.n_cells + 1), rng[:, 1] - rng[:, 0] + 1)
    for k in range(len(cf)):
        o1, o2 = own[cf[k] - 1]
        assert (cell_of[k], cn[k], cs[k]) in ((o1, o2, 1), (o2, o1, -1))


@pytest.mark.parametrize("name", ALL)
def test_closed_cells_and_volumes(meshes, name):
    """Sum over a cell's faces of nsign*A*n (+ boundary A*n) vanishes, the
    volumes are positive and add up to the domain volume."""
    m = meshes(name)
    n = m.n_cells
    S = (m.face_area[:, None] * m.face_normal.reshape(-1, 3))
    acc = np.zeros((n, 3))
    rng = m.cell_faces_range.reshape(-1, 2)
    cell_of = np.repeat(np.arange(n), rng[:, 1] - rng[:, 0] + 1)
    np.add.at(acc, cell_of, S[m.cell_faces - 1] * m.cell_nsign[:, None])
    np.add.at(acc, m.boundary_cellsID - 1, S[:m.n_bfaces])
    scale = np.sqrt(m.face_area.max())
    if m.dim == 2:
        scale = m.face_area.max()
    assert np.abs(acc).max() < 1e-10 * scale
    assert np.all(m.cell_volume > 0)
    pts = m.node_coords.reshape(-1, 3)
    ext = pts.max(0) - pts.min(0)
    if name in ("cavity40", "laplace3", "quad40", "box12", "box12_natural", "cavity_hex"):
        vol = ext[0] * ext[1] * (ext[2] if m.dim == 3 else 1.0)
        assert abs(m.cell_volume.sum() - vol) < 1e-10 * vol


def test_box_jitter_and_permutation(x):
    M = x.mesh
    a = M.box3d(6, 5, 4, jitter=0.1, seed_jitter=11, seed_perm=12)
    b = M.box3d(6, 5, 4, jitter=0.1, seed_jitter=11, seed_perm=12)
    c = M.box3d(6, 5, 4, jitter=0.1, seed_jitter=11, seed_perm=13)
    nat = M.box3d(6, 5, 4, jitter=0.0, seed_perm=0)
    assert np.array_equal(a.cell_centre, b.cell_centre)            # deterministic
    assert not np.array_equal(a.cell_faces, c.cell_faces)            # renumbered by the seed
    # same cells up to renumbering: volume multisets of unjittered boxes agree
    p = M.box3d(6, 5, 4, jitter=0.0, seed_perm=99)
    assert np.allclose(np.sort(p.cell_volume), np.sort(nat.cell_volume))
    h = np.array([1 / 6, 1 / 5, 1 / 4])
    d = np.abs(a.node_coords.reshape(-1, 3) - nat.node_coords.reshape(-1, 3))
    assert np.all(d <= 0.1 * h + 1e-15)
    # boundary faces keep exact axis-aligned normals
    bn = a.face_normal.reshape(-1, 3)[:a.n_bfaces]
    assert np.all(np.sort(np.abs(bn), axis=1)[:, :2] == 0.0)
    assert [a.n_cells, a.n_faces] == [120, 3 * 120 + 6 * 5 + 6 * 4 + 5 * 4]


def test_cavity_block_mesh_uniform(meshes):
    m = meshes("cavity40")
    assert np.allclose(m.cell_volume, 1 / 1600, rtol=1e-12)
    assert np.allclose(m.face_weight[m.n_bfaces:], 0.5, rtol=1e-12)
    assert np.allclose(m.face_delta[m.n_bfaces:], 1 / 40, rtol=1e-12)
    assert m.patch_names == ["bottom", "top", "left", "right"]


def test_unknown_patch_name(meshes):
    with pytest.raises(KeyError, match="is not a recognised boundary name"):
        meshes("cavity40").patch_index("inlet")
