"""ctypes front end of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of XCALibre's FP64 SIMPLE hot path that the
gfx950 library is checked against (and the timed CPU baseline).  Only tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py use it;
the product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"
_lib = None

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)


def build(force: bool = False) -> Path:
    """Compile liboracle.so from oracle.cpp (g++, -ffp-contract=off, OpenMP)."""
    src = HERE / "oracle.cpp"
    inc = HERE.parent / "include"
    if force or not SO.exists() or SO.stat().st_mtime < max(
            src.stat().st_mtime, (HERE / "oracle.h").stat().st_mtime,
            (inc / "xcal.h").stat().st_mtime):
        tmp = SO.with_suffix(".so.tmp")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                        "-std=c++17", f"-I{inc}", f"-I{HERE}", str(src), "-o", str(tmp)],
                       check=True)
        os.replace(tmp, SO)
    return SO


class _Term(C.Structure):
    _fields_ = [("type", C.c_int32), ("sign", C.c_int32), ("flux", _pd),
                ("flux_const", C.c_double)]


class _Src(C.Structure):
    _fields_ = [("x", _pd), ("y", _pd), ("z", _pd), ("sign", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        L.orc_last_error.restype = C.c_char_p
        L.orc_mesh_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_mesh_destroy.restype = None
        L.orc_nnz.argtypes = [C.c_void_p]
        L.orc_nnz.restype = C.c_int64
        L.orc_csr.argtypes = [C.c_void_p, _pi, _pi]
        L.orc_csr.restype = None
        L.orc_discretise_scalar.argtypes = [C.c_void_p, C.POINTER(_Term), C.c_int32,
                                            C.POINTER(_Src), C.c_int32, _pd, C.c_double, _pd, _pd]
        L.orc_discretise_vector.argtypes = [C.c_void_p, C.POINTER(_Term), C.c_int32,
                                            C.POINTER(_Src), C.c_int32, _pd, _pd, _pd,
                                            C.c_double, _pd, _pd, _pd, _pd]
        L.orc_apply_bcs.argtypes = [C.c_void_p, C.POINTER(_Term), C.c_int32, C.c_void_p,
                                    C.c_int32, C.c_int32, _pd, _pd, _pd, _pd, _pd]
        L.orc_relax_implicit_diagdom.argtypes = [C.c_void_p, _pd, _pd, _pd, C.c_double]
        L.orc_relax_implicit_diagdom.restype = None
        L.orc_set_reference.argtypes = [C.c_void_p, _pd, _pd, C.c_double, C.c_int64]
        L.orc_set_reference.restype = None
        L.orc_spmv.argtypes = [C.c_void_p, _pd, _pd, _pd]
        L.orc_spmv.restype = None
        L.orc_residual.argtypes = [C.c_void_p, _pd, _pd, _pd]
        L.orc_residual.restype = C.c_double
        L.orc_dilu_diagonal.argtypes = [C.c_void_p, _pd, _pd]
        L.orc_dilu_diagonal.restype = None
        L.orc_dilu_apply.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd]
        L.orc_dilu_apply.restype = None
        L.orc_krylov.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _pd, _pd, _pd, C.c_int32,
                                 C.c_double, C.c_double, _pi32]
        L.orc_interpolate.argtypes = [C.c_void_p, _pd, _pd]
        L.orc_interpolate.restype = None
        L.orc_correct_boundaries.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, _pd, _pd, _pd,
                                             _pd, _pd, _pd]
        L.orc_green_gauss.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd]
        L.orc_green_gauss.restype = None
        L.orc_green_gauss_tensor.argtypes = [C.c_void_p, _pd, _pd, _pd, C.POINTER(_pd)]
        L.orc_green_gauss_tensor.restype = None
        L.orc_grad_midpoint.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                        C.POINTER(_pd), C.POINTER(_pd), C.POINTER(_pd)]
        L.orc_div.argtypes = [C.c_void_p, _pd, _pd]
        L.orc_div.restype = None
        L.orc_flux.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd]
        L.orc_flux.restype = None
        L.orc_H.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]
        L.orc_H.restype = None
        L.orc_correct_mass_flux.argtypes = [C.c_void_p, _pd, _pd, _pd, C.c_int32, C.c_void_p,
                                            C.c_int32, _pd]
        L.orc_simple_run.argtypes = [C.c_void_p, C.c_void_p, _pd, _pd, _pd, _pd, C.c_int32, _pd,
                                     _pi32, _pi32, _pd]
        L.orc_simple_run_turb.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _pd, _pd, _pd, _pd,
                                          _pd, _pd, _pd, C.c_int32, _pd, _pi32, _pi32, _pd]
        L.orc_piso_run.argtypes = [C.c_void_p, C.c_void_p, _pd, _pd, _pd, _pd, C.c_int32,
                                   C.c_int32, _pd, _pi32, _pi32, _pd]
        L.orc_nonorth_correction.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd, _pd, _pd]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_smooth_jacobi.argtypes = [C.c_void_p, _pd, _pd, _pd, C.c_int32, C.c_double]
        L.orc_solve_system.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, _pd, _pd, _pd,
                                       C.POINTER(C.c_int32)]
        L.orc_kw_production.argtypes = [C.c_void_p, C.POINTER(_pd), _pd, C.c_double, _pd, _pd]
        L.orc_kw_wall_production.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, _pd,
                                             _pd, _pd, _pd, _pd]
        L.orc_kw_coeffs.argtypes = [C.c_void_p, _pd, _pd, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, _pd, _pd, _pd, _pd]
        L.orc_kw_constrain.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, _pd, _pd,
                                       _pd]
        L.orc_bound.argtypes = [C.c_void_p, _pd]
        L.orc_kw_nut.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, _pd, _pd, _pd,
                                 _pd]
        L.orc_kw_wall_constants.argtypes = [C.c_double, _pd]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_pd)


def _chk(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


class OracleMesh:
    """Oracle view of a host mesh (mesh.Mesh)."""

    def __init__(self, mesh, abi):
        self.mesh = mesh
        self.abi = abi
        self._host = abi.mesh_host(mesh)
        h = C.c_void_p()
        _chk(lib().orc_mesh_create(C.byref(self._host), C.byref(h)))
        self.h = h
        self.n = mesh.n_cells
        self.nnz = lib().orc_nnz(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_destroy(self.h)
            self.h = None

    def csr(self):
        rp = np.zeros(self.n + 1, np.int64)
        cv = np.zeros(self.nnz, np.int64)
        lib().orc_csr(self.h, rp.ctypes.data_as(_pi), cv.ctypes.data_as(_pi))
        return rp, cv

    @staticmethod
    def _terms(terms):
        arr = (_Term * len(terms))()
        keep = []
        for i, (typ, sign, flux) in enumerate(terms):
            if isinstance(flux, np.ndarray):
                f = np.ascontiguousarray(flux, dtype=np.float64)
                keep.append(f)
                arr[i] = _Term(typ, sign, _p(f), 0.0)
            else:
                arr[i] = _Term(typ, sign, None, float(flux))
        return arr, keep

    @staticmethod
    def _srcs(srcs):
        arr = (_Src * max(1, len(srcs)))()
        for i, (fields, sign) in enumerate(srcs):
            fs = fields if isinstance(fields, (list, tuple)) else [fields]
            ptrs = [_p(f) for f in fs] + [None] * (3 - len(fs))
            arr[i] = _Src(ptrs[0], ptrs[1], ptrs[2], sign)
        return arr

    def discretise_scalar(self, terms, srcs, prev=None, dt=1.0):
        t, keep = self._terms(terms)
        s = self._srcs(srcs)
        prev = np.zeros(self.n) if prev is None else prev
        nz = np.zeros(self.nnz)
        b = np.zeros(self.n)
        _chk(lib().orc_discretise_scalar(self.h, t, len(terms), s, len(srcs), _p(prev), dt,
                                         _p(nz), _p(b)))
        return nz, b

    def discretise_vector(self, terms, srcs, prev, dt=1.0):
        t, keep = self._terms(terms)
        s = self._srcs(srcs)
        nz = np.zeros(self.nnz)
        bx, by, bz = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        _chk(lib().orc_discretise_vector(self.h, t, len(terms), s, len(srcs), _p(prev[0]),
                                         _p(prev[1]), _p(prev[2]), dt, _p(nz), _p(bx), _p(by),
                                         _p(bz)))
        return nz, (bx, by, bz)

    def apply_bcs(self, terms, bcs, component, phi, nzval, b):
        """component: -1 for scalar equations (phi = array), 0..2 for vectors
        (phi = (x, y, z)).  nzval and b are updated in place."""
        t, keep = self._terms(terms)
        if component < 0:
            px, py, pz = phi, None, None
        else:
            px, py, pz = phi
        _chk(lib().orc_apply_bcs(self.h, t, len(terms), bcs, len(bcs), component, _p(px), _p(py),
                                 _p(pz), _p(nzval), _p(b)))

    def relax_implicit_diagdom(self, nzval, b, field, alpha):
        lib().orc_relax_implicit_diagdom(self.h, _p(nzval), _p(b), _p(field), alpha)

    def set_reference(self, nzval, b, pref, cell=0):
        lib().orc_set_reference(self.h, _p(nzval), _p(b), pref, cell)

    def spmv(self, nzval, x):
        y = np.zeros(self.n)
        lib().orc_spmv(self.h, _p(nzval), _p(np.ascontiguousarray(x)), _p(y))
        return y

    def residual(self, nzval, b, x):
        return lib().orc_residual(self.h, _p(nzval), _p(b), _p(x))

    def dilu_diagonal(self, nzval):
        D = np.zeros(self.n)
        lib().orc_dilu_diagonal(self.h, _p(nzval), _p(D))
        return D

    def dilu_apply(self, nzval, D, r):
        z = np.zeros(self.n)
        lib().orc_dilu_apply(self.h, _p(nzval), _p(D), _p(np.ascontiguousarray(r)), _p(z))
        return z

    def krylov(self, solver, precon, nzval, b, x0, itmax=1000, atol=None, rtol=0.1):
        atol = np.finfo(np.float64).eps ** 0.9 if atol is None else atol
        x = np.array(x0, dtype=np.float64, copy=True)
        it = C.c_int32()
        rc = _chk(lib().orc_krylov(self.h, solver, precon, _p(nzval), _p(b), _p(x), itmax, atol,
                                   rtol, C.byref(it)))
        return x, it.value, rc

    def smooth_jacobi(self, nzval, b, x0, loops, omega):
        """apply_smoother!(JacobiSmoother(loops, omega)) on a copy of x0."""
        x = np.array(x0, dtype=np.float64, copy=True)
        lib().orc_smooth_jacobi(self.h, _p(nzval), _p(np.ascontiguousarray(b, dtype=np.float64)),
                                _p(x), int(loops), float(omega))
        return x

    def solve_system(self, setup_c, nzval, b, x0, cn=False):
        """solve_system!: smoother, Krylov solve, Crank-Nicolson step (cn)."""
        x = np.array(x0, dtype=np.float64, copy=True)
        it = C.c_int32()
        rc = _chk(lib().orc_solve_system(self.h, C.byref(setup_c), int(bool(cn)), _p(nzval),
                                         _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x),
                                         C.byref(it)))
        return x, it.value, rc

    # ---- k-omega pieces
    def kw_production(self, gU, nut, alpha1):
        g = [np.ascontiguousarray(gU[i], dtype=np.float64) for i in range(9)]
        arr = (_pd * 9)(*[_p(a) for a in g])
        Pk, Pw = np.zeros(self.n), np.zeros(self.n)
        lib().orc_kw_production(self.h, arr, _p(np.ascontiguousarray(nut, dtype=np.float64)),
                                float(alpha1), _p(Pk), _p(Pw))
        return Pk, Pw

    def kw_wall_production(self, bc_k, nu, U, k, Pk):
        bc = self.abi.bc_array(bc_k)
        out = np.array(Pk, dtype=np.float64, copy=True)
        U = [np.ascontiguousarray(U[i], dtype=np.float64) for i in range(3)]
        _chk(lib().orc_kw_wall_production(self.h, bc, len(bc_k), float(nu), _p(U[0]), _p(U[1]),
                                          _p(U[2]), _p(np.ascontiguousarray(k, dtype=np.float64)),
                                          _p(out)))
        return out

    def kw_coeffs(self, omega, nutf, nu, beta1, beta_star, sigma_omega, sigma_k):
        Dw, Dk = np.zeros(self.n), np.zeros(self.n)
        mw, mk = np.zeros(self.mesh.n_faces), np.zeros(self.mesh.n_faces)
        lib().orc_kw_coeffs(self.h, _p(np.ascontiguousarray(omega, dtype=np.float64)),
                            _p(np.ascontiguousarray(nutf, dtype=np.float64)), float(nu),
                            float(beta1), float(beta_star), float(sigma_omega), float(sigma_k),
                            _p(Dw), _p(Dk), _p(mw), _p(mk))
        return Dw, Dk, mw, mk

    def kw_constrain(self, bc_omega, nu, k, nzval, b):
        bc = self.abi.bc_array(bc_omega)
        A = np.array(nzval, dtype=np.float64, copy=True)
        bb = np.array(b, dtype=np.float64, copy=True)
        _chk(lib().orc_kw_constrain(self.h, bc, len(bc_omega), float(nu),
                                    _p(np.ascontiguousarray(k, dtype=np.float64)), _p(A), _p(bb)))
        return A, bb

    def bound(self, v):
        out = np.array(v, dtype=np.float64, copy=True)
        lib().orc_bound(self.h, _p(out))
        return out

    def kw_nut(self, bc_nut, nu, k, omega):
        bc = self.abi.bc_array(bc_nut)
        nut, nutf = np.zeros(self.n), np.zeros(self.mesh.n_faces)
        _chk(lib().orc_kw_nut(self.h, bc, len(bc_nut), float(nu),
                              _p(np.ascontiguousarray(k, dtype=np.float64)),
                              _p(np.ascontiguousarray(omega, dtype=np.float64)), _p(nut),
                              _p(nutf)))
        return nut, nutf

    @staticmethod
    def kw_wall_constants(nu):
        out = np.zeros(4)
        lib().orc_kw_wall_constants(float(nu), _p(out))
        return dict(cmu25=out[0], yplus_lam=out[1], kappa=out[2], E=out[3])

    def interpolate(self, phi):
        f = np.zeros(self.mesh.n_faces)
        lib().orc_interpolate(self.h, _p(phi), _p(f))
        return f

    def correct_boundaries(self, bcs, phi, phif):
        if isinstance(phi, (list, tuple)):
            _chk(lib().orc_correct_boundaries(self.h, bcs, len(bcs), _p(phi[0]), _p(phi[1]),
                                              _p(phi[2]), _p(phif[0]), _p(phif[1]), _p(phif[2])))
        else:
            _chk(lib().orc_correct_boundaries(self.h, bcs, len(bcs), _p(phi), None, None,
                                              _p(phif), None, None))

    def green_gauss(self, phif):
        g = [np.zeros(self.n) for _ in range(3)]
        lib().orc_green_gauss(self.h, _p(phif), _p(g[0]), _p(g[1]), _p(g[2]))
        return g

    def green_gauss_tensor(self, fx, fy, fz):
        T = [np.zeros(self.n) for _ in range(9)]
        ptrs = (_pd * 9)(*[_p(t) for t in T])
        lib().orc_green_gauss_tensor(self.h, _p(fx), _p(fy), _p(fz), ptrs)
        return T

    def grad_midpoint(self, bcs, phi):
        """grad! with Grad{Midpoint}: phi a cell array (scalar) or three
        (vector); returns (face values, gradient components)."""
        ph = [np.ascontiguousarray(phi, dtype=np.float64)] if np.ndim(phi[0]) == 0 else \
            [np.ascontiguousarray(c, dtype=np.float64) for c in phi]
        nc = len(ph)
        f = [np.zeros(self.mesh.n_faces) for _ in range(nc)]
        g = [np.zeros(self.n) for _ in range(3 * nc)]
        arr = lambda xs: (_pd * len(xs))(*[_p(v) for v in xs])
        _chk(lib().orc_grad_midpoint(self.h, bcs, len(bcs), nc, arr(ph), arr(f), arr(g)))
        return f, g

    def div(self, psif):
        d = np.zeros(self.n)
        lib().orc_div(self.h, _p(psif), _p(d))
        return d

    def flux(self, fx, fy, fz):
        m = np.zeros(self.mesh.n_faces)
        lib().orc_flux(self.h, _p(fx), _p(fy), _p(fz), _p(m))
        return m

    def H(self, nzval, b, U):
        H = [np.zeros(self.n) for _ in range(3)]
        lib().orc_H(self.h, _p(nzval), _p(b[0]), _p(b[1]), _p(b[2]), _p(U[0]), _p(U[1]),
                    _p(U[2]), _p(H[0]), _p(H[1]), _p(H[2]))
        return H

    def correct_mass_flux(self, p_nzval, p, pflux, psign, bcs, mdot):
        _chk(lib().orc_correct_mass_flux(self.h, _p(p_nzval), _p(p), _p(pflux), psign, bcs,
                                         len(bcs), _p(mdot)))

    def nonorth_correction(self, grad, p, flux, b):
        """nonorthogonal_face_correction: b (modified copy returned) plus the
        non-orthogonal face corrections of gradient `grad` (3, n) and field p
        with face coefficients `flux`."""
        b = np.array(b, dtype=np.float64, copy=True)
        g = [np.ascontiguousarray(grad[c], dtype=np.float64) for c in range(3)]
        _chk(lib().orc_nonorth_correction(self.h, _p(g[0]), _p(g[1]), _p(g[2]),
                                          _p(np.ascontiguousarray(p, dtype=np.float64)),
                                          _p(np.ascontiguousarray(flux, dtype=np.float64)),
                                          _p(b)))
        return b

    def piso_run(self, setup_c, U, p, steps, inner_loops=2):
        """Transient PISO (Solvers_2_PISO.jl:40-211): `steps` time steps of
        setup_c.dt with `inner_loops` pressure correctors each."""
        Ux, Uy, Uz = (np.array(U[i], dtype=np.float64, copy=True) for i in range(3))
        pp = np.array(p, dtype=np.float64, copy=True)
        res = np.zeros(4 * steps)
        kit = np.zeros(4 * steps, np.int32)
        done = C.c_int32()
        mdot = np.zeros(self.mesh.n_faces)
        _chk(lib().orc_piso_run(self.h, C.byref(setup_c), _p(Ux), _p(Uy), _p(Uz), _p(pp), steps,
                                inner_loops, _p(res), C.byref(done), kit.ctypes.data_as(_pi32),
                                _p(mdot)))
        d = done.value
        return dict(U=np.stack([Ux, Uy, Uz]), p=pp, mdotf=mdot, done=d,
                    residuals=res.reshape(4, steps)[:, :d],
                    krylov_iters=kit.reshape(4, steps)[:, :d])

    def simple_run(self, setup_c, U, p, iterations, turb=None, k=None, omega=None, nut=None):
        """setup_c: _abi.SimpleSetupC; U: (3, n) array (modified copy returned).
        turb: _abi.TurbulenceSetupC for the k-omega model, with initial k, omega, nut."""
        Ux, Uy, Uz = (np.array(U[i], dtype=np.float64, copy=True) for i in range(3))
        pp = np.array(p, dtype=np.float64, copy=True)
        res = np.zeros(4 * iterations)
        kit = np.zeros(4 * iterations, np.int32)
        done = C.c_int32()
        mdot = np.zeros(self.mesh.n_faces)
        out = {}
        if turb is None:
            _chk(lib().orc_simple_run(self.h, C.byref(setup_c), _p(Ux), _p(Uy), _p(Uz), _p(pp),
                                      iterations, _p(res), C.byref(done),
                                      kit.ctypes.data_as(_pi32), _p(mdot)))
        else:
            kk, ww, nn = (np.array(v, dtype=np.float64, copy=True) for v in (k, omega, nut))
            _chk(lib().orc_simple_run_turb(self.h, C.byref(setup_c), C.byref(turb), _p(Ux),
                                           _p(Uy), _p(Uz), _p(pp), _p(kk), _p(ww), _p(nn),
                                           iterations, _p(res), C.byref(done),
                                           kit.ctypes.data_as(_pi32), _p(mdot)))
            out = dict(k=kk, omega=ww, nut=nn)
        d = done.value
        return dict(U=np.stack([Ux, Uy, Uz]), p=pp, mdotf=mdot, done=d,
                    residuals=res.reshape(4, iterations)[:, :d],
                    krylov_iters=kit.reshape(4, iterations)[:, :d], **out)
