"""Result output in the reference's VTK formats (src/IOFormats/VTK/).

``initialise_writer(VTK(), mesh)`` and ``write_results(iteration, time, mesh,
meshData, BCs, args...; suffix)`` have the reference's names and argument
meanings (VTK_writer.jl:4-6, VTK_writer_3D.jl:16,150); the files are written
by ``libxcal_mesh.so`` (include/xcal_io.h) from host copies of the fields.
A field is a numpy array of one value per cell: shape (n,) for a scalar,
(3, n) for a vector, (9, n) or (3, 3, n) for a tensor -- the layouts
``api.Physics`` and ``hip.Simple.get`` use.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _libs


@dataclass(frozen=True)
class VTK:
    """Output format selector VTK() (IOFormats/IOFormats.jl)."""


class _Writer:
    def __init__(self, handle: C.c_void_p, dim: int, n_cells: int):
        self._h = handle
        self.dim = dim
        self.n_cells = n_cells

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _libs.mesh_lib().xcalw_free(h)
            self._h = None

    def name(self, iteration: int, time, suffix: str = "") -> str:
        buf = C.create_string_buffer(512)
        n = _libs.mesh_lib().xcalw_result_name(self._h, int(iteration), float(time),
                                                suffix.encode(), buf, 512)
        if n < 0:
            raise ValueError("result file name too long")
        return buf.value.decode()


class VTKWriter2D(_Writer):
    """VTKWriter2D (VTK_types.jl): legacy ASCII .vtk for Mesh2."""


class VTKWriter3D(_Writer):
    """VTKWriter3D (VTK_types.jl): XML .vtu of polyhedra for Mesh3."""


def _check(rc: int):
    if rc != 0:
        raise RuntimeError(_libs.mesh_lib().xcalw_last_error().decode())


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def initialise_writer(format, mesh):
    """initialise_writer(format, mesh): VTK_writer.jl:4 (2D),
    VTK_writer_3D.jl:16-146 (3D, builds the polyhedron header once)."""
    if not isinstance(format, VTK):
        raise TypeError("the writers here produce VTK() output")
    lib = _libs.mesh_lib()
    i64 = {k: np.ascontiguousarray(getattr(mesh, k), dtype=np.int64)
           for k in ("cell_nodes_range", "cell_nodes", "face_nodes_range", "face_nodes",
                     "face_owner")}
    f64 = {k: np.ascontiguousarray(getattr(mesh, k), dtype=np.float64)
           for k in ("node_coords", "cell_centre", "face_centre")}
    h = C.c_void_p()
    if mesh.dim == 2:
        _check(lib.xcalw_vtk2d_init(mesh.n_nodes, _p(f64["node_coords"], C.c_double),
                                    mesh.n_cells, _p(i64["cell_nodes_range"], C.c_int64),
                                    _p(i64["cell_nodes"], C.c_int64), C.byref(h)))
        return VTKWriter2D(h, 2, mesh.n_cells)
    _check(lib.xcalw_vtu3d_init(mesh.n_nodes, _p(f64["node_coords"], C.c_double), mesh.n_cells,
                                _p(i64["cell_nodes_range"], C.c_int64),
                                _p(i64["cell_nodes"], C.c_int64),
                                _p(f64["cell_centre"], C.c_double), mesh.n_faces,
                                _p(i64["face_nodes_range"], C.c_int64),
                                _p(i64["face_nodes"], C.c_int64),
                                _p(i64["face_owner"], C.c_int64),
                                _p(f64["face_centre"], C.c_double), C.byref(h)))
    return VTKWriter3D(h, 3, mesh.n_cells)


class _FieldC(C.Structure):
    _fields_ = [("label", C.c_char_p), ("ncomp", C.c_int32), ("values", C.POINTER(C.c_double))]


def _field(label: str, values, n: int):
    a = np.asarray(values, dtype=np.float64)
    if a.ndim == 3 and a.shape[:2] == (3, 3):
        a = a.reshape(9, -1)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.ndim != 2 or a.shape[0] not in (1, 3, 9) or a.shape[1] != n:
        raise TypeError('Input data should be a ScalarField, VectorField or TensorField '
                        f'e.g. ("U", U): got {label!r} with shape {np.shape(values)}')
    return np.ascontiguousarray(a)


def write_results(iteration: int, time, mesh, meshData: _Writer, BCs, *args, suffix: str = "",
                  directory=None):
    """write_results(iteration, time, mesh, meshData, BCs, args...; suffix)
    (VTK_writer.jl:6, VTK_writer_3D.jl:150).  args are (label, field)
    pairs.  The file goes to `directory` (default: the working directory, as
    in the reference).  Returns the path written."""
    del BCs  # unused by the VTK writers, as in the reference
    if meshData.n_cells != mesh.n_cells:
        raise ValueError("writer and mesh differ")
    keep = [(_field(lbl, v, mesh.n_cells), lbl.encode()) for lbl, v in args]
    arr = (_FieldC * max(1, len(keep)))()
    for i, (a, lbl) in enumerate(keep):
        arr[i] = _FieldC(lbl, a.shape[0], _p(a, C.c_double))
    d = None if directory is None else str(directory).encode()
    _check(_libs.mesh_lib().xcalw_write_results(meshData._h, int(iteration), float(time),
                                                 suffix.encode(), d, len(keep), arr))
    name = meshData.name(iteration, time, suffix)
    return name if directory is None else f"{directory}/{name}"


def format_double(x: float) -> str:
    """The text Julia's print(::Float64) gives for x (used for every value
    in the files)."""
    buf = C.create_string_buffer(40)
    n = _libs.mesh_lib().xcalw_format_double(float(x), buf)
    return buf.raw[:n].decode()
