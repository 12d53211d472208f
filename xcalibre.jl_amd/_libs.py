"""Loading of the in-tree native libraries.

``libxcal_mesh.so`` (host mesh builders and result writers) and ``libxcalhip.so`` (the gfx950
HIP library behind the C ABI in include/xcal.h) are built in-tree by
``build.py``.  Nothing here falls back to Python or CPU code: a missing
library is an error.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
MESH_SO = PKG_DIR / "libxcal_mesh.so"
HIP_SO = PKG_DIR / "libxcalhip.so"

_mesh = None
_hip = None

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u64 = C.c_uint64
_dbl = C.c_double
_pdbl = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)
_cp = C.c_char_p


def _need(path: Path) -> None:
    if not path.exists():
        raise RuntimeError(
            f"{path.name} is not built; run __graft_entry__.build() (or python "
            f"xcalibre.jl_amd/build.py) first")


def mesh_lib():
    global _mesh
    if _mesh is None:
        _need(MESH_SO)
        lib = C.CDLL(str(MESH_SO))
        lib.xcalm_last_error.restype = _cp
        lib.xcalm_read_foam.argtypes = [_cp, _dbl, C.POINTER(_vp)]
        lib.xcalm_read_unv2.argtypes = [_cp, _dbl, C.POINTER(_vp)]
        lib.xcalm_box3d.argtypes = [_i64, _i64, _i64, _dbl, _dbl, _dbl, _dbl, _u64, _u64,
                                    C.POINTER(_vp)]
        lib.xcalm_block2d.argtypes = [_i64, _pdbl, _i64, _pi64, _i64, _pi64, _i32,
                                      C.POINTER(_cp), _pi64, _pi64, C.POINTER(_vp)]
        lib.xcalm_array.argtypes = [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]
        lib.xcalm_sizes.argtypes = [_vp, C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_i64),
                                    C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i32)]
        lib.xcalm_patch_name.argtypes = [_vp, _i32]
        lib.xcalm_patch_name.restype = _cp
        lib.xcalm_free.argtypes = [_vp]
        lib.xcalm_free.restype = None
        lib.xcalw_last_error.restype = _cp
        lib.xcalw_format_double.argtypes = [_dbl, _cp]
        lib.xcalw_format_double.restype = _i32
        lib.xcalw_vtk2d_init.argtypes = [_i64, _pdbl, _i64, _pi64, _pi64, C.POINTER(_vp)]
        lib.xcalw_vtu3d_init.argtypes = [_i64, _pdbl, _i64, _pi64, _pi64, _pdbl, _i64, _pi64,
                                         _pi64, _pi64, _pdbl, C.POINTER(_vp)]
        lib.xcalw_write_results.argtypes = [_vp, _i64, _dbl, _cp, _cp, _i32, _vp]
        lib.xcalw_result_name.argtypes = [_vp, _i64, _dbl, _cp, _cp, _i32]
        lib.xcalw_result_name.restype = _i32
        lib.xcalw_foam_init.argtypes = [_cp, _i32, _i64, _pdbl, _i64, _i64, _i64, _pi64, _pi64,
                                        _pi64, _i32, C.POINTER(_cp), _pi64, C.POINTER(_vp)]
        lib.xcalw_foam_write_results.argtypes = [_vp, _i64, _dbl, _i32, _vp]
        lib.xcalw_foam_time_dir.argtypes = [_i64, _dbl, _cp, _i32]
        lib.xcalw_foam_time_dir.restype = _i32
        lib.xcalw_free.argtypes = [_vp]
        lib.xcalw_free.restype = None
        _mesh = lib
    return _mesh
