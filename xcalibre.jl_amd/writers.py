"""Result output in the reference's formats (src/IOFormats/VTK/, OpenFOAM/).

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


@dataclass(frozen=True)
class OpenFOAM:
    """Output format selector OpenFOAM() (IOFormats/OpenFOAM/OpenFOAM_types.jl)."""


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


class FOAMWriter(_Writer):
    """FOAMWriter (OpenFOAM_types.jl): an OpenFOAM case directory whose
    constant/polyMesh was written by initialise_writer."""

    def __init__(self, handle, n_cells, directory):
        super().__init__(handle, 3, n_cells)
        self.directory = directory

    def name(self, iteration: int, time, suffix=None) -> str:
        buf = C.create_string_buffer(64)
        n = _libs.mesh_lib().xcalw_foam_time_dir(int(iteration), float(time), buf, 64)
        if n < 0:
            raise ValueError("time directory name too long")
        return buf.value.decode()


class _FoamBC(C.Structure):
    _fields_ = [("patch", C.c_int32), ("kind", C.c_int32), ("value", C.c_double * 3)]


class _FoamFieldC(C.Structure):
    _fields_ = [("label", C.c_char_p), ("ncomp", C.c_int32), ("values", C.POINTER(C.c_double)),
                ("n_bcs", C.c_int32), ("bcs", C.POINTER(_FoamBC))]


def _check(rc: int):
    if rc != 0:
        raise RuntimeError(_libs.mesh_lib().xcalw_last_error().decode())


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def initialise_writer(format, mesh, directory=None):
    """initialise_writer(format, mesh): VTK_writer.jl:4 (2D),
    VTK_writer_3D.jl:16-146 (3D, builds the polyhedron header once),
    OpenFOAM_writer.jl:1-189 (writes XCALibre.foam and constant/polyMesh
    under `directory`, default the working directory)."""
    if isinstance(format, OpenFOAM):
        return _foam_writer(mesh, directory)
    if not isinstance(format, VTK):
        raise TypeError("output must be VTK() or OpenFOAM()")
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


def _foam_writer(mesh, directory):
    lib = _libs.mesh_lib()
    i64 = {k: np.ascontiguousarray(getattr(mesh, k), dtype=np.int64)
           for k in ("face_nodes_range", "face_nodes", "face_owner", "patch_range")}
    X = np.ascontiguousarray(mesh.node_coords, dtype=np.float64)
    names = [n.encode() for n in mesh.patch_names]
    cnames = (C.c_char_p * max(1, len(names)))(*names)
    h = C.c_void_p()
    d = None if directory is None else str(directory).encode()
    _check(lib.xcalw_foam_init(d, mesh.dim, mesh.n_nodes, _p(X, C.c_double), mesh.n_cells,
                               mesh.n_faces, mesh.n_bfaces, _p(i64["face_nodes_range"], C.c_int64),
                               _p(i64["face_nodes"], C.c_int64), _p(i64["face_owner"], C.c_int64),
                               len(names), cnames, _p(i64["patch_range"], C.c_int64),
                               C.byref(h)))
    return FOAMWriter(h, mesh.n_cells, directory)


def _write_foam(iteration, time, mesh, w: FOAMWriter, BCs, args, suffix):
    """OpenFOAM_writer.jl:191-273: each field's boundaryField lists the BCs of
    the field of the same label (suffix None), or of p / U for a scalar /
    vector postprocessing field (suffix "")."""
    keep, fields = [], []
    for label, values in args:
        a = _field(label, values, mesh.n_cells)
        if a.shape[0] == 9:
            raise TypeError('Input data should be a ScalarField or VectorField e.g. ("U", U)')
        key = label if suffix is None else ("p" if a.shape[0] == 1 else "U")
        bcs = (_FoamBC * max(1, len(BCs[key])))()
        for i, (kind, patch, value) in enumerate(BCs[key]):
            v = np.broadcast_to(np.asarray(value, dtype=np.float64), (3,))
            bcs[i] = _FoamBC(patch, kind, (C.c_double * 3)(*v))
        lbl = label.encode()
        keep += [a, bcs, lbl]
        fields.append(_FoamFieldC(lbl, a.shape[0], _p(a, C.c_double), len(BCs[key]), bcs))
    arr = (_FoamFieldC * max(1, len(fields)))(*fields)
    _check(_libs.mesh_lib().xcalw_foam_write_results(w._h, int(iteration), float(time),
                                                      len(fields), arr))
    name = w.name(iteration, time)
    return name if w.directory is None else f"{w.directory}/{name}"


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


def write_results(iteration: int, time, mesh, meshData: _Writer, BCs, *args, suffix=None,
                  directory=None):
    """write_results(iteration, time, mesh, meshData, BCs, args...; suffix)
    (VTK_writer.jl:6, VTK_writer_3D.jl:150, OpenFOAM_writer.jl:191).  args
    are (label, field) pairs; BCs is ``Configuration.boundaries`` (read by the
    OpenFOAM writer only, as in the reference).  A VTK file goes to
    `directory` (default: the working directory, as in the reference); an
    OpenFOAM time directory goes into the writer's case directory.  Returns
    the path written."""
    if meshData.n_cells != mesh.n_cells:
        raise ValueError("writer and mesh differ")
    if isinstance(meshData, FOAMWriter):
        return _write_foam(iteration, time, mesh, meshData, BCs, args, suffix)
    suffix = "" if suffix is None else suffix
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
