"""ctypes bindings of libxcalhip.so (include/xcal.h).

Thin wrappers only: every call goes to the gfx950 library; a missing library
or a failing call raises.  Handles free their device memory when collected.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from . import _libs

_vp = C.c_void_p
_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)
_lib = None


class XcalError(RuntimeError):
    pass


class ItmaxWarning(UserWarning):
    pass


def lib():
    global _lib
    if _lib is None:
        _libs._need(_libs.HIP_SO)
        L = C.CDLL(str(_libs.HIP_SO))
        L.xcal_last_error.restype = C.c_char_p
        L.xcal_abi_version.restype = C.c_int32
        sig = {
            "xcal_ctx_create": [C.c_int32, C.POINTER(_vp)],
            "xcal_ctx_destroy": [_vp],
            "xcal_ctx_synchronize": [_vp],
            "xcal_mesh_create": [_vp, C.POINTER(A.MeshHost), C.POINTER(_vp)],
            "xcal_mesh_destroy": [_vp],
            "xcal_mesh_info": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _pi32, _pi32,
                               C.POINTER(C.c_int64)],
            "xcal_mesh_csr_pattern": [_vp, _pi32, _pi32],
            "xcal_connectivity_create": [_vp, C.POINTER(A.MeshHost), C.c_int32, C.POINTER(_vp)],
            "xcal_connectivity_info": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _pi32,
                                       _pi32, _pd, _pd],
            "xcal_connectivity_get": [_vp, C.c_int32, _pi32],
            "xcal_connectivity_destroy": [_vp],
            "xcal_mesh_create_batch": [_vp, C.POINTER(C.POINTER(A.MeshHost)), C.c_int32,
                                       C.POINTER(_vp)],
            "xcal_mesh_cases": [_vp, _pi32, C.POINTER(C.c_int64)],
            "xcal_mesh_sweep_form": [_vp, _pi32, _pi32],
            "xcal_mesh_sweep_far": [_vp, _pi32, _pi32],
            "xcal_mesh_pipe_tuning": [_vp, _pi32, _pi32, _pd],
            "xcal_mesh_cone_form": [_vp, _pi32, _pi32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
            "xcal_mesh_cone_bands": [_vp, _pi32, _pi32],
            "xcal_solve_system_batch": [_vp, C.POINTER(A.SolverSetupC), _vp, C.c_int32,
                                        C.POINTER(A.SolveStats)],
            "xcal_simple_run_batch": [_vp, C.c_int32, _pd, _pi32, _pi32],
            "xcal_simple_set_case_nu": [_vp, _pd],
            "xcal_simple_set_turbulence": [_vp, C.POINTER(A.TurbulenceSetupC)],
            "xcal_server_create": [C.POINTER(A.ServerConfigC), C.POINTER(_vp)],
            "xcal_server_destroy": [_vp],
            "xcal_server_submit": [_vp, C.POINTER(A.MeshHost), C.POINTER(A.SimpleSetupC),
                                   C.POINTER(A.TurbulenceSetupC), _pd, _pd, _pd, _pd, _pd,
                                   C.POINTER(C.c_int64)],
            "xcal_server_poll": [_vp, C.c_int64, C.POINTER(A.RequestStatusC)],
            "xcal_server_wait": [_vp, C.c_int64, C.c_double],
            "xcal_server_fetch": [_vp, C.c_int64, C.c_int32, _pd],
            "xcal_server_release": [_vp, C.c_int64],
            "xcal_field_create": [_vp, C.c_int32, C.POINTER(_vp)],
            "xcal_field_destroy": [_vp],
            "xcal_field_upload": [_vp, C.c_int32, _pd],
            "xcal_field_download": [_vp, C.c_int32, _pd],
            "xcal_field_fill": [_vp, C.c_int32, C.c_double],
            "xcal_eqn_create": [_vp, C.c_int32, C.POINTER(_vp)],
            "xcal_eqn_destroy": [_vp],
            "xcal_discretise": [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp],
            "xcal_discretise_dt": [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp, C.c_double],
            "xcal_update_equation": [_vp],
            "xcal_apply_bcs": [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32],
            "xcal_relax_implicit_diagdom": [_vp, _vp, C.c_int32, C.c_double],
            "xcal_set_reference": [_vp, C.c_double, C.c_int64],
            "xcal_eqn_download": [_vp, C.c_int32, C.c_int32, _pd, _pd],
            "xcal_precond_update": [_vp, C.c_int32],
            "xcal_precond_apply": [_vp, _vp, C.c_int32, _vp, C.c_int32],
            "xcal_precond_apply_noalloc": [_vp, _vp, C.c_int32, _vp, C.c_int32],
            "xcal_spmv": [_vp, _vp, C.c_int32, _vp, C.c_int32],
            "xcal_solve_system": [_vp, C.POINTER(A.SolverSetupC), _vp, C.c_int32,
                                  C.POINTER(A.SolveStats)],
            "xcal_interpolate": [_vp, _vp],
            "xcal_correct_boundaries": [_vp, _vp, _vp, C.c_int32],
            "xcal_green_gauss": [_vp, _vp],
            "xcal_grad": [_vp, _vp, _vp, _vp, C.c_int32],
            "xcal_grad_midpoint": [_vp, _vp, _vp, _vp, C.c_int32],
            "xcal_div": [_vp, _vp],
            "xcal_flux": [_vp, _vp],
            "xcal_relax_explicit": [_vp, _vp, C.c_double],
            "xcal_inverse_diagonal": [_vp, _vp],
            "xcal_remove_pressure_source": [_vp, _vp, C.c_int32],
            "xcal_H": [_vp, _vp, _vp],
            "xcal_correct_mass_flux": [_vp, _vp, _vp, _vp, C.c_int32, _vp, C.c_int32],
            "xcal_correct_velocity": [_vp, _vp, _vp, _vp],
            "xcal_nonorthogonal_correction": [_vp, _vp, _vp, _vp],
            "xcal_simple_create": [_vp, C.POINTER(A.SimpleSetupC), C.POINTER(_vp)],
            "xcal_simple_destroy": [_vp],
            "xcal_simple_set_field": [_vp, C.c_int32, _pd],
            "xcal_simple_get_field": [_vp, C.c_int32, _pd],
            "xcal_simple_run": [_vp, C.c_int32, _pd, _pi32, _pi32],
            "xcal_simple_converged": [_vp, _pi32],
            "xcal_piso_run": [_vp, C.c_int32, C.c_int32, _pd, _pi32, _pi32],
            "xcal_simple_stage_times": [_vp, _pd],
            "xcal_kernel_timing_enable": [_vp, C.c_int32, C.c_int32],
            "xcal_kernel_timing_read": [_vp, C.c_int32, C.POINTER(C.c_int64), _pd],
        }
        for name, argt in sig.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = C.c_int
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise XcalError(lib().xcal_last_error().decode())
    return rc


def _pd_of(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pd)


class Context:
    """One HIP stream on one device (xcal_ctx)."""

    def __init__(self, device: int = 0):
        self.h = _vp()
        check(lib().xcal_ctx_create(device, C.byref(self.h)))

    def synchronize(self):
        check(lib().xcal_ctx_synchronize(self.h))

    KERNELS = {"cg_spmv": 0, "dilu_fwd": 1, "dilu_bwd": 2}

    def time_kernel(self, name: str, enable: bool = True):
        """Bracket every 8th launch of a pressure-solve kernel with HIP events
        (kernels are enabled independently)."""
        check(lib().xcal_kernel_timing_enable(self.h, self.KERNELS[name], int(enable)))

    def kernel_time(self, name: str):
        """(timed launches, total milliseconds) of a kernel since time_kernel()."""
        n, ms = C.c_int64(), C.c_double()
        check(lib().xcal_kernel_timing_read(self.h, self.KERNELS[name], C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_ctx_destroy(self.h)
            self.h = None


class Connectivity:
    """xcal_connectivity: a mesh's CSR pattern, slot maps and DILU levels
    built on the device (on_device=True) or by the host loops."""

    NAMES = ("rowptr", "colval", "diag", "cf_slot", "trans", "face_slot", "flev", "blev")

    def __init__(self, ctx: Context | None, mesh, on_device: bool = True):
        self.h = _vp()
        host = A.mesh_host(mesh)
        check(lib().xcal_connectivity_create(ctx.h if ctx is not None else None, C.byref(host),
                                             int(on_device), C.byref(self.h)))
        nnz, ncf_, fl, bl = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        wall, dev = C.c_double(), C.c_double()
        check(lib().xcal_connectivity_info(self.h, C.byref(nnz), C.byref(ncf_), C.byref(fl),
                                           C.byref(bl), C.byref(wall), C.byref(dev)))
        self.nnz, self.levels = nnz.value, (fl.value, bl.value)
        self.wall_ms, self.device_ms = wall.value, dev.value
        n, nf, ncf = mesh.n_cells, mesh.n_faces, ncf_.value
        sizes = (n + 1, self.nnz, n, ncf, self.nnz, nf, n, n)
        self.arrays = {}
        for w, (name, size) in enumerate(zip(self.NAMES, sizes)):
            a = np.zeros(size, np.int32)
            check(lib().xcal_connectivity_get(self.h, w, a.ctypes.data_as(_pi32)))
            self.arrays[name] = a

    def __getitem__(self, name):
        return self.arrays[name]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_connectivity_destroy(self.h)
            self.h = None


class BatchMeshInfo:
    """Host view of a batch: the cases' meshes and their concatenated sizes."""

    def __init__(self, meshes):
        self.cases = list(meshes)
        self.n_cells = sum(m.n_cells for m in self.cases)
        self.n_faces = sum(m.n_faces for m in self.cases)
        self.patch_names = list(self.cases[0].patch_names)
        self.cell_offsets = np.cumsum([0] + [m.n_cells for m in self.cases])
        self.face_offsets = np.cumsum([0] + [m.n_faces for m in self.cases])


class DeviceMesh:
    """xcal_mesh.  DeviceMesh(ctx, mesh) uploads one mesh; DeviceMesh.batch(ctx,
    meshes) uploads independent cases as one batch mesh (xcal_mesh_create_batch)."""

    def __init__(self, ctx: Context, mesh, _batch=None):
        self.ctx = ctx
        self.mesh = mesh
        self.h = _vp()
        if _batch is None:
            host = A.mesh_host(mesh)
            check(lib().xcal_mesh_create(ctx.h, C.byref(host), C.byref(self.h)))
            self.ncase = 1
        else:
            hosts = [A.mesh_host(m) for m in _batch]
            arr = (C.POINTER(A.MeshHost) * len(hosts))(*[C.pointer(h) for h in hosts])
            check(lib().xcal_mesh_create_batch(ctx.h, arr, len(hosts), C.byref(self.h)))
            self.ncase = len(hosts)
        n, nnz, bytes_ = C.c_int64(), C.c_int64(), C.c_int64()
        fl, bl = C.c_int32(), C.c_int32()
        check(lib().xcal_mesh_info(self.h, C.byref(n), C.byref(nnz), C.byref(fl), C.byref(bl),
                                   C.byref(bytes_)))
        self.n, self.nnz = n.value, nnz.value
        self.levels = (fl.value, bl.value)
        self.device_bytes = bytes_.value

    @classmethod
    def batch(cls, ctx: Context, meshes):
        return cls(ctx, BatchMeshInfo(meshes), _batch=list(meshes))

    def sweep_form(self):
        """(forward, backward) entry bounds of the pipelined batch sweeps, -1
        where the sweeps step level by level."""
        f, b = C.c_int32(), C.c_int32()
        check(lib().xcal_mesh_sweep_form(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def sweep_far(self):
        """(forward, backward): whether the pipelined batch sweeps read
        dependencies beyond their LDS ring from global memory."""
        f, b = C.c_int32(), C.c_int32()
        check(lib().xcal_mesh_sweep_far(self.h, C.byref(f), C.byref(b)))
        return bool(f.value), bool(b.value)

    def pipe_tuning(self):
        """(forward rows per thread, backward rows per thread, {(direction,
        rows): µs per launch}) of the pipelined batch sweeps as tuned."""
        f, b = C.c_int32(), C.c_int32()
        us = np.zeros(6)
        check(lib().xcal_mesh_pipe_tuning(self.h, C.byref(f), C.byref(b), _pd_of(us)))
        t = {(d, r): float(us[3 * i + j]) for i, d in enumerate(("fwd", "bwd"))
             for j, r in enumerate((1, 2, 4))}
        return f.value, b.value, t

    def cone_bands(self):
        """(forward, backward) bands of levels, one launch each, of the cone
        sweeps; (0, 0) where the sweeps launch per level."""
        f, b = C.c_int32(), C.c_int32()
        check(lib().xcal_mesh_cone_bands(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def cone_form(self):
        """(rows per block, forward blocks, forward cone positions, backward
        cone positions) of the cone sweeps; all 0 where the sweeps launch per
        level."""
        r, nb = C.c_int32(), C.c_int32()
        f, b = C.c_int64(), C.c_int64()
        check(lib().xcal_mesh_cone_form(self.h, C.byref(r), C.byref(nb), C.byref(f), C.byref(b)))
        return r.value, nb.value, f.value, b.value

    def csr_pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        cv = np.zeros(self.nnz, np.int32)
        check(lib().xcal_mesh_csr_pattern(self.h, rp.ctypes.data_as(_pi32), cv.ctypes.data_as(_pi32)))
        return rp, cv

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_mesh_destroy(self.h)
            self.h = None


_NCOMP = {A.CELL_SCALAR: 1, A.CELL_VECTOR: 3, A.FACE_SCALAR: 1, A.FACE_VECTOR: 3,
          A.CELL_TENSOR: 9}


class Field:
    def __init__(self, dmesh: DeviceMesh, kind: int):
        self.dmesh = dmesh
        self.kind = kind
        self.ncomp = _NCOMP[kind]
        self.len = dmesh.mesh.n_faces if kind in (A.FACE_SCALAR, A.FACE_VECTOR) else dmesh.n
        self.h = _vp()
        check(lib().xcal_field_create(dmesh.h, kind, C.byref(self.h)))

    def upload(self, values):
        v = np.asarray(values, dtype=np.float64).reshape(self.ncomp, self.len)
        for c in range(self.ncomp):
            check(lib().xcal_field_upload(self.h, c, _pd_of(np.ascontiguousarray(v[c]))))
        return self

    def download(self) -> np.ndarray:
        out = np.zeros((self.ncomp, self.len))
        for c in range(self.ncomp):
            row = np.zeros(self.len)
            check(lib().xcal_field_download(self.h, c, _pd_of(row)))
            out[c] = row
        return out[0] if self.ncomp == 1 else out

    def fill(self, value, component: int = 0):
        check(lib().xcal_field_fill(self.h, component, float(value)))
        return self

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_field_destroy(self.h)
            self.h = None


class _TermC(C.Structure):
    _fields_ = [("type", C.c_int32), ("sign", C.c_int32), ("flux", _vp), ("flux_const", C.c_double)]


class _SrcC(C.Structure):
    _fields_ = [("field", _vp), ("sign", C.c_int32)]


def terms_c(terms):
    """terms: list of (type, sign, Field-or-constant)."""
    arr = (_TermC * len(terms))()
    for i, (typ, sign, flux) in enumerate(terms):
        if isinstance(flux, Field):
            arr[i] = _TermC(typ, sign, flux.h, 0.0)
        else:
            arr[i] = _TermC(typ, sign, None, float(flux))
    return arr


def sources_c(srcs):
    arr = (_SrcC * max(1, len(srcs)))()
    for i, (field, sign) in enumerate(srcs):
        arr[i] = _SrcC(field.h, sign)
    return arr


class Equation:
    """ScalarEquation / VectorEquation storage on the device."""

    def __init__(self, dmesh: DeviceMesh, vector: bool):
        self.dmesh = dmesh
        self.vector = vector
        self.h = _vp()
        check(lib().xcal_eqn_create(dmesh.h, int(vector), C.byref(self.h)))

    def discretise(self, terms, srcs, prev=None, dt=None):
        """discretise!(eqn, prev, config); dt = config.runtime.dt for Time{Euler}."""
        ph = prev.h if prev is not None else None
        if dt is None:
            check(lib().xcal_discretise(self.h, terms_c(terms), len(terms), sources_c(srcs),
                                        len(srcs), ph))
        else:
            check(lib().xcal_discretise_dt(self.h, terms_c(terms), len(terms), sources_c(srcs),
                                           len(srcs), ph, float(dt)))

    def update_equation(self):
        check(lib().xcal_update_equation(self.h))

    def apply_bcs(self, terms, bcs, phi, component=0):
        check(lib().xcal_apply_bcs(self.h, terms_c(terms), len(terms), bcs, len(bcs), phi.h,
                                   component))

    def relax_implicit_diagdom(self, phi, component, alpha):
        check(lib().xcal_relax_implicit_diagdom(self.h, phi.h, component, alpha))

    def set_reference(self, pref, cell=1):
        check(lib().xcal_set_reference(self.h, pref, cell))

    def download(self, which=0, component=0):
        nz = np.zeros(self.dmesh.nnz)
        b = np.zeros(self.dmesh.n)
        check(lib().xcal_eqn_download(self.h, which, component, _pd_of(nz), _pd_of(b)))
        return nz, b

    def precond_update(self, precon):
        check(lib().xcal_precond_update(self.h, precon))

    def precond_apply(self, r: Field, z: Field, rc=0, zc=0):
        check(lib().xcal_precond_apply(self.h, r.h, rc, z.h, zc))

    def precond_apply_noalloc(self, r: Field, z: Field, rc=0, zc=0):
        """precond_apply staged in the library's fixed device array (no
        allocation; meshes of at most 2**20 rows)."""
        check(lib().xcal_precond_apply_noalloc(self.h, r.h, rc, z.h, zc))

    def spmv(self, xf: Field, y: Field, xc=0, yc=0):
        check(lib().xcal_spmv(self.h, xf.h, xc, y.h, yc))

    def solve(self, setup: A.SolverSetupC, phi: Field, component=0):
        st = A.SolveStats()
        rc = check(lib().xcal_solve_system(self.h, C.byref(setup), phi.h, component, C.byref(st)))
        self.last_nouter = st.nouter  # refinement sweeps (1 for FP64)
        return st.niter, st.residual, rc

    def solve_batch(self, setup: A.SolverSetupC, phi: Field, component=0):
        """Per-case (niter, nouter, residual) arrays and the status of a batch solve."""
        K = self.dmesh.ncase
        st = (A.SolveStats * K)()
        rc = check(lib().xcal_solve_system_batch(self.h, C.byref(setup), phi.h, component, st))
        return (np.array([s.niter for s in st]), np.array([s.nouter for s in st]),
                np.array([s.residual for s in st]), rc)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_eqn_destroy(self.h)
            self.h = None


class Simple:
    """Device-resident SIMPLE run (xcal_simple_*)."""

    def __init__(self, dmesh: DeviceMesh, setup: A.SimpleSetupC):
        self.dmesh = dmesh
        self._setup = setup  # keeps BC arrays alive
        self.h = _vp()
        check(lib().xcal_simple_create(dmesh.h, C.byref(setup), C.byref(self.h)))

    def set_U(self, U):
        U = np.ascontiguousarray(np.asarray(U, dtype=np.float64).reshape(3, self.dmesh.n))
        check(lib().xcal_simple_set_field(self.h, 0, _pd_of(U)))

    def set_p(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(self.dmesh.n))
        check(lib().xcal_simple_set_field(self.h, 1, _pd_of(p)))

    def set_turbulence(self, tsetup: A.TurbulenceSetupC):
        self._tsetup = tsetup  # keeps BC arrays alive
        check(lib().xcal_simple_set_turbulence(self.h, C.byref(tsetup)))

    def set_scalar(self, which: str, values):
        sel = {"p": 1, "k": 3, "omega": 4, "nut": 5}[which]
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(self.dmesh.n))
        check(lib().xcal_simple_set_field(self.h, sel, _pd_of(v)))

    def get(self, which: str) -> np.ndarray:
        sel = {"U": 0, "p": 1, "mdotf": 2, "k": 3, "omega": 4, "nut": 5}[which]
        n = self.dmesh.mesh.n_faces if which == "mdotf" else self.dmesh.n
        out = np.zeros((3 if which == "U" else 1) * n)
        check(lib().xcal_simple_get_field(self.h, sel, _pd_of(out)))
        return out.reshape(3, n) if which == "U" else out

    def run(self, iterations: int):
        res = np.zeros(4 * iterations)
        kit = np.zeros(4 * iterations, np.int32)
        done = C.c_int32()
        check(lib().xcal_simple_run(self.h, iterations, _pd_of(res), C.byref(done),
                                    kit.ctypes.data_as(_pi32)))
        d = done.value
        conv = C.c_int32()
        check(lib().xcal_simple_converged(self.h, C.byref(conv)))
        return dict(done=d, residuals=res.reshape(4, iterations)[:, :d],
                    krylov_iters=kit.reshape(4, iterations)[:, :d], converged=bool(conv.value))

    def run_piso(self, steps: int, inner_loops: int = 2):
        """Transient PISO time steps (setup dt); residuals per step."""
        res = np.zeros(4 * steps)
        kit = np.zeros(4 * steps, np.int32)
        done = C.c_int32()
        check(lib().xcal_piso_run(self.h, steps, inner_loops, _pd_of(res), C.byref(done),
                                  kit.ctypes.data_as(_pi32)))
        d = done.value
        return dict(done=d, residuals=res.reshape(4, steps)[:, :d],
                    krylov_iters=kit.reshape(4, steps)[:, :d])

    def run_batch(self, iterations: int):
        """Batch meshes: residuals and Krylov counts per case, (ncase, 4, done)."""
        K = self.dmesh.ncase
        res = np.zeros(K * 4 * iterations)
        kit = np.zeros(K * 4 * iterations, np.int32)
        done = C.c_int32()
        check(lib().xcal_simple_run_batch(self.h, iterations, _pd_of(res), C.byref(done),
                                          kit.ctypes.data_as(_pi32)))
        d = done.value
        return dict(done=d, residuals=res.reshape(K, 4, iterations)[:, :, :d],
                    krylov_iters=kit.reshape(K, 4, iterations)[:, :, :d])

    def set_case_nu(self, nu):
        nu = np.ascontiguousarray(np.asarray(nu, dtype=np.float64).reshape(self.dmesh.ncase))
        check(lib().xcal_simple_set_case_nu(self.h, _pd_of(nu)))

    def stage_times(self):
        t = np.zeros(6)
        check(lib().xcal_simple_stage_times(self.h, _pd_of(t)))
        return dict(momentum_assembly=t[0], momentum_solve=t[1], pressure_assembly=t[2],
                    pressure_solve=t[3], other=t[4], total=t[5])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_simple_destroy(self.h)
            self.h = None


class Server:
    """Serving engine (xcal_server_*): continuous batching of independent
    steady SIMPLE cases; at most `slots` cases in flight, stepped together in
    shared launches, under an HBM budget.  `lanes` batches step concurrently,
    each on its own stream; a case joins the lane closest to it in size."""

    def __init__(self, device=0, slots=4, hbm_budget=64 << 30, max_iterations=1000, lanes=1):
        cfg = A.ServerConfigC(device, slots, int(hbm_budget), max_iterations, lanes)
        self.h = _vp()
        check(lib().xcal_server_create(C.byref(cfg), C.byref(self.h)))
        self._cells = {}

    def submit(self, mesh, setup: A.SimpleSetupC, U0=None, p0=None, turbulence=None, k0=None,
               omega0=None, nut0=None) -> int:
        """turbulence: None (laminar) or an A.TurbulenceSetupC; the initial
        fields may be None (zero start)."""
        host = A.mesh_host(mesh)
        arrs = [None if v is None else np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1))
                for v in (U0, p0, k0, omega0, nut0)]
        rid = C.c_int64()
        check(lib().xcal_server_submit(self.h, C.byref(host), C.byref(setup),
                                       None if turbulence is None else C.byref(turbulence),
                                       *[None if a is None else _pd_of(a) for a in arrs],
                                       C.byref(rid)))
        self._cells[rid.value] = (mesh.n_cells, turbulence is not None
                                  and turbulence.model != A.TURB_LAMINAR)
        return rid.value

    def status(self, rid: int) -> dict:
        st = A.RequestStatusC()
        check(lib().xcal_server_poll(self.h, rid, C.byref(st)))
        return dict(state=st.state, iterations=st.iterations, cells=st.cells,
                    footprint=st.footprint, t_submit=st.t_submit, t_start=st.t_start,
                    t_end=st.t_end, converged=bool(st.converged), lane=st.lane,
                    batch_rebuilds=st.batch_rebuilds)

    def wait(self, rid: int = 0, timeout: float = 0.0) -> bool:
        """True when finished, False on timeout."""
        return check(lib().xcal_server_wait(self.h, rid, float(timeout))) == 0

    def fetch(self, rid: int) -> dict:
        """Fields of a finished request: U (3, n), p, and k, omega, nut for
        turbulent cases."""
        n, turb = self._cells[rid]
        out = {}
        names = [("U", A.SIMPLE_U), ("p", A.SIMPLE_P)]
        if turb:
            names += [("k", A.SIMPLE_K), ("omega", A.SIMPLE_OMEGA), ("nut", A.SIMPLE_NUT)]
        for name, w in names:
            v = np.zeros((3 if name == "U" else 1) * n)
            check(lib().xcal_server_fetch(self.h, rid, w, _pd_of(v)))
            out[name] = v.reshape(3, n) if name == "U" else v
        return out

    def release(self, rid: int):
        check(lib().xcal_server_release(self.h, rid))
        self._cells.pop(rid, None)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.xcal_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
