"""Host meshes in XCALibre's Mesh2/Mesh3 array layout.

A :class:`Mesh` holds numpy copies of the arrays that ``libxcal_mesh.so``
builds (1-based Int64 indices, Float64 geometry, boundary faces first), the
same arrays ``adapt(backend, mesh)`` would expose in the reference
(src/Mesh/Mesh_0_types.jl:135-185).  The builders mirror the reference's
entry points:

* :func:`FOAM3D_mesh`  -- src/FoamMesh/FoamMesh_5_build.jl:21-27
* :func:`UNV2D_mesh`   -- src/UNV2/UNV2_3_builder.jl:21-49
* :func:`block_mesh_2d`-- BlockMesher2D ``generate!`` + ``update_mesh_format``
  (src/Mesh/BlockMesher2D/6_generate.jl:3-23, src/UNV2/UNV2_4_update_format.jl)
* :func:`box3d`        -- seeded jittered, renumbered hex box routed through the
  FOAM path (the generator of the 1M-cell workload).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _libs

_ARRAYS = (
    "cell_centre", "cell_volume", "cell_faces_range", "cell_nodes_range",
    "cell_nodes", "cell_faces", "cell_neighbours", "cell_nsign",
    "face_nodes_range", "face_nodes", "face_owner", "face_centre",
    "face_normal", "face_e", "face_area", "face_delta", "face_weight",
    "boundary_cellsID", "patch_range", "node_coords",
)


@dataclass
class Mesh:
    dim: int
    n_cells: int
    n_faces: int
    n_bfaces: int
    n_nodes: int
    patch_names: list
    arrays: dict = field(repr=False)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    @property
    def n_ifaces(self) -> int:
        return self.n_faces - self.n_bfaces

    def patch_index(self, name: str) -> int:
        """0-based patch index; error text as Discretise_4_assign_boundaries.jl:40."""
        try:
            return self.patch_names.index(name)
        except ValueError:
            raise KeyError(f'"{name}" is not a recognised boundary name') from None

    def __str__(self):  # Mesh_0_types.jl:187-201
        kind = "2D" if self.dim == 2 else "3D"
        return (f"{kind} Mesh with:\n-> {self.n_cells} cells\n-> {self.n_faces} faces\n"
                f"-> {self.n_nodes} nodes")


def _wrap(handle: C.c_void_p) -> Mesh:
    lib = _libs.mesh_lib()
    dim = C.c_int32()
    nc, nf, nbf, nn = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    npatch = C.c_int32()
    lib.xcalm_sizes(handle, C.byref(dim), C.byref(nc), C.byref(nf), C.byref(nbf),
                    C.byref(nn), C.byref(npatch))
    arrays = {}
    for name in _ARRAYS:
        ptr, ln, isf = C.c_void_p(), C.c_int64(), C.c_int32()
        rc = lib.xcalm_array(handle, name.encode(), C.byref(ptr), C.byref(ln), C.byref(isf))
        if rc != 0:
            raise RuntimeError(lib.xcalm_last_error().decode())
        dt = np.float64 if isf.value else np.int64
        if ln.value == 0:
            arrays[name] = np.zeros(0, dtype=dt)
        else:
            buf = (C.c_double if isf.value else C.c_int64) * ln.value
            arrays[name] = np.frombuffer(buf.from_address(ptr.value), dtype=dt).copy()
    names = [lib.xcalm_patch_name(handle, i).decode() for i in range(npatch.value)]
    lib.xcalm_free(handle)
    return Mesh(dim.value, nc.value, nf.value, nbf.value, nn.value, names, arrays)


def _check(rc: int):
    if rc != 0:
        raise RuntimeError(_libs.mesh_lib().xcalm_last_error().decode())


def FOAM3D_mesh(polymesh_dir, scale: float = 1.0) -> Mesh:
    h = C.c_void_p()
    _check(_libs.mesh_lib().xcalm_read_foam(str(polymesh_dir).encode(), C.c_double(scale),
                                            C.byref(h)))
    return _wrap(h)


def UNV2D_mesh(path, scale: float = 1.0) -> Mesh:
    h = C.c_void_p()
    _check(_libs.mesh_lib().xcalm_read_unv2(str(path).encode(), C.c_double(scale), C.byref(h)))
    return _wrap(h)


def box3d(nx: int, ny: int, nz: int, L=(1.0, 1.0, 1.0), jitter: float = 0.0,
          seed_jitter: int = 0x5EED0002, seed_perm: int = 0x5EED0003) -> Mesh:
    """Hex box, interior nodes jittered by +-jitter*h, cells renumbered by a
    seeded Fisher-Yates permutation (seed_perm=0 keeps natural order)."""
    h = C.c_void_p()
    _check(_libs.mesh_lib().xcalm_box3d(nx, ny, nz, C.c_double(L[0]), C.c_double(L[1]),
                                        C.c_double(L[2]), C.c_double(jitter),
                                        C.c_uint64(seed_jitter), C.c_uint64(seed_perm),
                                        C.byref(h)))
    return _wrap(h)


def block_mesh_2d(points, edges, blocks, patches) -> Mesh:
    """BlockMesher2D.

    points  : list of (x, y, z) corner points (``Point``)
    edges   : list of (p1, p2, ncells), 1-based point ids, in ``line!`` call order
    blocks  : list of 4-tuples of 1-based edge ids as ``quad(edges, [...])``
    patches : list of (name, [edge ids]) as ``Patch(name, [...])``
    """
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    eds = np.ascontiguousarray(np.asarray(edges, dtype=np.int64).reshape(-1, 3))
    bls = np.ascontiguousarray(np.asarray(blocks, dtype=np.int64).reshape(-1, 4))
    names = [p[0] for p in patches]
    ptr = np.zeros(len(patches) + 1, dtype=np.int64)
    flat = []
    for i, (_, ids) in enumerate(patches):
        flat.extend(ids)
        ptr[i + 1] = len(flat)
    flat = np.asarray(flat, dtype=np.int64)
    cnames = (C.c_char_p * len(names))(*[n.encode() for n in names])
    h = C.c_void_p()
    i64p = C.POINTER(C.c_int64)
    _check(_libs.mesh_lib().xcalm_block2d(
        len(pts), pts.ctypes.data_as(C.POINTER(C.c_double)), len(eds), eds.ctypes.data_as(i64p),
        len(bls), bls.ctypes.data_as(i64p), len(names), cnames, ptr.ctypes.data_as(i64p),
        flat.ctypes.data_as(i64p), C.byref(h)))
    return _wrap(h)


def cavity2d(n: int, L: float = 1.0) -> Mesh:
    """Unit-square lid-driven cavity from one BlockMesher2D block with patches
    bottom, top, left, right (edges e1 south, e2 north, e3 west, e4 east)."""
    points = [(0.0, 0.0, 0.0), (L, 0.0, 0.0), (0.0, L, 0.0), (L, L, 0.0)]
    edges = [(1, 2, n), (3, 4, n), (1, 3, n), (2, 4, n)]
    return block_mesh_2d(points, edges, [(1, 2, 3, 4)],
                         [("bottom", [1]), ("top", [2]), ("left", [3]), ("right", [4])])


def backward_facing_step_2d(step: float, H: float = 1.0, Lin: float = 1.0, Lout: float = 4.0,
                            nx_in: int = 10, nx_out: int = 40, ny_low: int = 5,
                            ny_up: int = 5) -> Mesh:
    """Three-block backward-facing step: an inlet channel of height H-step
    above the step, and a downstream channel of height H (two blocks stacked).
    Patches: inlet, outlet, walls, top."""
    h = step
    p = [(0.0, h, 0.0), (Lin, h, 0.0), (Lin + Lout, h, 0.0),
         (0.0, H, 0.0), (Lin, H, 0.0), (Lin + Lout, H, 0.0),
         (Lin, 0.0, 0.0), (Lin + Lout, 0.0, 0.0)]
    edges = [
        (1, 2, nx_in),     # 1 inlet-block south (step top)
        (4, 5, nx_in),     # 2 inlet-block north
        (1, 4, ny_up),     # 3 inlet
        (2, 5, ny_up),     # 4 interface inlet-block / upper-out block
        (2, 3, nx_out),    # 5 upper-out south = lower-out north (interface)
        (5, 6, nx_out),    # 6 upper-out north
        (3, 6, ny_up),     # 7 outlet upper
        (7, 8, nx_out),    # 8 lower-out south (bottom wall)
        (7, 2, ny_low),    # 9 step face (west of lower-out)
        (8, 3, ny_low),    # 10 outlet lower
    ]
    blocks = [(1, 2, 3, 4), (5, 6, 4, 7), (8, 5, 9, 10)]
    patches = [("inlet", [3]), ("outlet", [7, 10]), ("walls", [1, 8, 9]), ("top", [2, 6])]
    return block_mesh_2d(p, edges, blocks, patches)
