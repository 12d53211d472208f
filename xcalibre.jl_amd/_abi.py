"""ctypes mirrors of the structs and constants in include/xcal.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

# operator types
TIME_STEADY, TIME_EULER, LAPLACIAN_LINEAR, DIV_LINEAR, DIV_UPWIND, DIV_LUST, \
    DIV_BOUNDED_UPWIND, SI = range(8)
TIME_CRANK_NICOLSON = 8
# boundary condition kinds
BC_DIRICHLET, BC_NEUMANN, BC_ZEROGRADIENT, BC_EXTRAPOLATED, BC_WALL, BC_SYMMETRY = range(6)
BC_KWALL, BC_OMEGAWALL, BC_NUTWALL, BC_EMPTY = 6, 7, 8, 9
# solvers / preconditioners / precision
SOLVER_CG, SOLVER_BICGSTAB = 0, 1
PRECON_NONE, PRECON_JACOBI, PRECON_DILU = 0, 1, 2
PREC_FP64, PREC_FP32_IR, PREC_FP32_PRECON = 0, 1, 2
# field kinds
CELL_SCALAR, CELL_VECTOR, FACE_SCALAR, FACE_VECTOR, CELL_TENSOR = range(5)
# status
OK, ITMAX = 0, 1

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int64)


class MeshHost(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("n_patches", C.c_int32),
        ("n_cells", C.c_int64), ("n_faces", C.c_int64), ("n_bfaces", C.c_int64),
        ("n_cell_faces", C.c_int64),
        ("cell_centre", _pd), ("cell_volume", _pd), ("cell_faces_range", _pi),
        ("cell_faces", _pi), ("cell_neighbours", _pi), ("cell_nsign", _pi),
        ("face_owner", _pi), ("face_centre", _pd), ("face_normal", _pd), ("face_e", _pd),
        ("face_area", _pd), ("face_delta", _pd), ("face_weight", _pd),
        ("boundary_cellsID", _pi), ("patch_range", _pi),
    ]


def mesh_host(mesh) -> MeshHost:
    """Borrowed view of a :class:`mesh.Mesh` as ``xcal_mesh_host``.  The numpy
    arrays stay owned by the mesh object, which must outlive the struct's use."""
    a = mesh.arrays

    def dp(name):
        arr = a[name]
        assert arr.dtype == np.float64 and arr.flags.c_contiguous
        return arr.ctypes.data_as(_pd)

    def ip(name):
        arr = a[name]
        assert arr.dtype == np.int64 and arr.flags.c_contiguous
        return arr.ctypes.data_as(_pi)

    return MeshHost(
        mesh.dim, len(mesh.patch_names), mesh.n_cells, mesh.n_faces, mesh.n_bfaces,
        len(a["cell_faces"]), dp("cell_centre"), dp("cell_volume"), ip("cell_faces_range"),
        ip("cell_faces"), ip("cell_neighbours"), ip("cell_nsign"), ip("face_owner"),
        dp("face_centre"), dp("face_normal"), dp("face_e"), dp("face_area"), dp("face_delta"),
        dp("face_weight"), ip("boundary_cellsID"), ip("patch_range"))


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("patch", C.c_int32), ("value", C.c_double * 3)]


class SolverSetupC(C.Structure):
    _fields_ = [("solver", C.c_int32), ("precon", C.c_int32), ("precision", C.c_int32),
                ("itmax", C.c_int32), ("convergence", C.c_double), ("relax", C.c_double),
                ("atol", C.c_double), ("rtol", C.c_double), ("ir_inner_rtol", C.c_double),
                ("smoother_loops", C.c_int32), ("smoother_omega", C.c_double)]


class SimpleSetupC(C.Structure):
    _fields_ = [("nu", C.c_double), ("div_scheme", C.c_int32), ("has_pref", C.c_int32),
                ("pref", C.c_double), ("U", SolverSetupC), ("p", SolverSetupC),
                ("bc_U", C.POINTER(BC)), ("bc_p", C.POINTER(BC)), ("n_bcs", C.c_int32),
                ("grad_p", C.c_int32), ("grad_U", C.c_int32), ("time_U", C.c_int32),
                ("dt", C.c_double), ("ncorrectors", C.c_int32)]


GRAD_GAUSS, GRAD_MIDPOINT = 0, 1


class ServerConfigC(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("hbm_budget", C.c_int64),
                ("max_iterations", C.c_int32), ("lanes", C.c_int32)]


class RequestStatusC(C.Structure):
    _fields_ = [("state", C.c_int32), ("iterations", C.c_int32), ("cells", C.c_int64),
                ("footprint", C.c_int64), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("converged", C.c_int32), ("lane", C.c_int32),
                ("batch_rebuilds", C.c_int64)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class TurbulenceSetupC(C.Structure):
    _fields_ = [("model", C.c_int32), ("beta_star", C.c_double), ("alpha1", C.c_double),
                ("beta1", C.c_double), ("sigma_k", C.c_double), ("sigma_omega", C.c_double),
                ("div_scheme", C.c_int32), ("k", SolverSetupC), ("omega", SolverSetupC),
                ("bc_k", C.POINTER(BC)), ("bc_omega", C.POINTER(BC)), ("bc_nut", C.POINTER(BC))]


TURB_LAMINAR, TURB_KOMEGA = 0, 1
SIMPLE_U, SIMPLE_P, SIMPLE_MDOTF, SIMPLE_K, SIMPLE_OMEGA, SIMPLE_NUT = 0, 1, 2, 3, 4, 5


class SolveStats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("nouter", C.c_int32), ("residual", C.c_double)]


def bc_array(bcs):
    """ctypes array of BC from a list of (kind, patch, value) tuples."""
    arr = (BC * len(bcs))()
    for i, (kind, patch, value) in enumerate(bcs):
        v = list(value) if np.ndim(value) else [float(value), 0.0, 0.0]
        v = (v + [0.0, 0.0, 0.0])[:3]
        arr[i] = BC(kind, patch, (C.c_double * 3)(*v))
    return arr
