"""Host-side mirror of XCALibre's user API for the SIMPLE path.

The names, keyword arguments, defaults and error behaviour follow the
reference so that a case script reads the same:

    mesh = block_mesh_2d(...)                      # or FOAM3D_mesh / UNV2D_mesh
    model = Physics(time=Steady(), fluid=Fluid(Incompressible, nu=0.01),
                    turbulence=RANS(Laminar), energy=Energy(Isothermal),
                    domain=mesh)
    BCs = assign(region=mesh, U=[Wall("top", [1, 0, 0]), ...], p=[...])
    solvers = dict(U=SolverSetup(solver=Bicgstab(), preconditioner=DILU(),
                                 convergence=1e-7, relax=0.7), p=...)
    schemes = dict(U=Schemes(divergence=Upwind), p=Schemes())
    config = Configuration(solvers=solvers, schemes=schemes,
                           runtime=Runtime(iterations=200, write_interval=-1,
                                           time_step=1),
                           hardware=Hardware(backend=HIPBackend(), workgroup=256),
                           boundaries=BCs)
    initialise(model.momentum.U, [0, 0, 0])
    residuals = run(model, config, pref=0.0)       # run!(model, config; pref)

References: Physics (ModelPhysics/0_type_definition.jl:196-229), assign
(Discretise_4_assign_boundaries.jl:5-49), SolverSetup/Runtime/Schemes
(Solve_1_api.jl:8-211), Configuration/Hardware (Simulate_0_types.jl:32-73),
run! (Solvers_3_solver_dispatch.jl:120-135), residual NamedTuple
(Solvers_1_SIMPLE.jl:262).  Execution always goes through libxcalhip.so.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _abi as A
from . import hip
from .writers import VTK, OpenFOAM, initialise_writer, write_results  # noqa: F401


# ------------------------------------------------------------- schemes
class SteadyState: ...
class Euler: ...
# Crank-Nicolson assembles as Euler (Discretise_1_schemes.jl:44-62); the
# equation whose first term it is then takes x = 2x - x_old after each solve
# (Solve_1_api.jl:291-296).
class CrankNicolson: ...


# operator type of each time scheme (xcal_term.type)
TIME = {SteadyState: A.TIME_STEADY, Euler: A.TIME_EULER,
        CrankNicolson: A.TIME_CRANK_NICOLSON}
class Linear: ...
class Upwind: ...
class LUST: ...
class BoundedUpwind: ...
class Gauss: ...
class Midpoint: ...


_DIV = {Linear: A.DIV_LINEAR, Upwind: A.DIV_UPWIND, LUST: A.DIV_LUST,
        BoundedUpwind: A.DIV_BOUNDED_UPWIND}
_GRAD = {Gauss: A.GRAD_GAUSS, Midpoint: A.GRAD_MIDPOINT}


@dataclass
class Schemes:
    time: Any = SteadyState
    divergence: Any = Linear
    laplacian: Any = Linear
    gradient: Any = Gauss
    limiter: Any = None


# ------------------------------------------------------- solver setup
class Cg: ...
class Bicgstab: ...
class Jacobi: ...
class DILU: ...


@dataclass
class JacobiSmoother:
    """Weighted Jacobi smoother (Smoothers_jacobi.jl:36-51): `loops` sweeps
    with weight `omega` on the field before each Krylov solve.
    JacobiSmoother(mesh) gives 5 loops and omega 1;
    JacobiSmoother(domain=mesh, loops=L, omega=2/3) the keyword form."""
    domain: Any = None
    loops: int = 5
    omega: float = 1.0

    def __init__(self, mesh=None, *, domain=None, loops=None, omega=2.0 / 3.0):
        if mesh is not None:  # JacobiSmoother(mesh): 5 loops, omega 1
            self.domain, self.loops, self.omega = mesh, 5, 1.0
            return
        if domain is None or loops is None:
            raise TypeError("JacobiSmoother(; domain, loops, omega=2/3) needs domain and loops")
        self.domain, self.loops, self.omega = domain, int(loops), float(omega)


@dataclass
class SolverSetup:
    solver: Any
    preconditioner: Any
    convergence: float
    relax: float
    float_type: Any = np.float64
    smoother: Any = None
    limit: Any = None
    itmax: int = 1000
    atol: Optional[float] = None
    rtol: float = 0.1
    # "fp64"; "fp32-ir" (FP32 inner Krylov, FP64 refinement); "fp32-precon"
    # (FP64 Krylov with the DILU sweeps in FP32, single meshes in cone form)
    precision: str = "fp64"
    inner_rtol: float = 1e-6     # fp32-ir: relative tolerance of each inner solve

    def __post_init__(self):
        if self.atol is None:
            self.atol = float(np.finfo(self.float_type).eps) ** 0.9
        if self.smoother is not None and not isinstance(self.smoother, JacobiSmoother):
            raise NotImplementedError(f"smoother {self.smoother!r} is not supported")

    def c(self) -> A.SolverSetupC:
        solver = {Cg: A.SOLVER_CG, Bicgstab: A.SOLVER_BICGSTAB}.get(type(self.solver))
        if solver is None:
            raise NotImplementedError(f"solver {self.solver!r} is not supported")
        pre = {Jacobi: A.PRECON_JACOBI, DILU: A.PRECON_DILU, type(None): A.PRECON_NONE}.get(
            type(self.preconditioner))
        if pre is None:
            raise NotImplementedError(f"preconditioner {self.preconditioner!r} is not supported")
        prec = {"fp64": A.PREC_FP64, "fp32-ir": A.PREC_FP32_IR,
                "fp32-precon": A.PREC_FP32_PRECON}[self.precision]
        sm = self.smoother
        return A.SolverSetupC(solver, pre, prec, int(self.itmax), float(self.convergence),
                              float(self.relax), float(self.atol), float(self.rtol),
                              float(self.inner_rtol), int(sm.loops) if sm else 0,
                              float(sm.omega) if sm else 0.0)


@dataclass
class Runtime:
    iterations: int
    write_interval: int
    time_step: float = 1.0


class HIPBackend:
    """The gfx950 device backend (sibling of CPU()/ROCBackend())."""

    def __init__(self, device: int = 0):
        self.device = device


@dataclass
class Hardware:
    backend: Any
    workgroup: int = 256


@dataclass
class Configuration:
    solvers: dict
    schemes: dict
    runtime: Runtime
    hardware: Hardware
    boundaries: dict
    postprocess: Any = None


# ------------------------------------------------------------- physics
class Steady: ...
class Transient: ...  # run! dispatches to PISO (Solvers_3_solver_dispatch.jl:186-190)
class Incompressible: ...
class Laminar: ...
class Isothermal: ...


@dataclass
class Fluid:
    kind: Any = Incompressible
    nu: float = 0.0


@dataclass
class KOmega:
    """RANS{KOmega} coefficients (RANS_kOmega.jl:41-45); k, omega, nut are
    cell fields allocated by Physics."""
    beta_star: float = 0.09
    alpha1: float = 0.52
    beta1: float = 0.072
    sigma_k: float = 0.5
    sigma_omega: float = 0.5
    k: np.ndarray = None
    omega: np.ndarray = None
    nut: np.ndarray = None


def RANS(kind=Laminar, **coeffs):
    """RANS{Laminar}() or RANS{KOmega}(; β⁺, α1, β1, σk, σω)."""
    if kind is Laminar:
        return Laminar()
    if kind is KOmega:
        names = {"β⁺": "beta_star", "α1": "alpha1", "β1": "beta1", "σk": "sigma_k",
                 "σω": "sigma_omega"}
        return KOmega(**{names.get(k, k): float(v) for k, v in coeffs.items()})
    raise NotImplementedError(f"RANS{{{kind}}} is not on the device path")


def Energy(kind=Isothermal):
    return kind()


@dataclass
class Momentum:
    U: np.ndarray
    p: np.ndarray


@dataclass
class Physics:
    domain: Any
    fluid: Fluid
    time: Any = field(default_factory=Steady)
    turbulence: Any = field(default_factory=Laminar)
    energy: Any = field(default_factory=Isothermal)
    momentum: Momentum = None

    def __post_init__(self):
        n = self.domain.n_cells
        self.momentum = Momentum(np.zeros((3, n)), np.zeros(n))
        if isinstance(self.turbulence, KOmega):
            t = self.turbulence
            t.k, t.omega, t.nut = np.zeros(n), np.zeros(n), np.zeros(n)


def initialise(field_values: np.ndarray, value):
    """initialise!(field, value) (Fields_0_types.jl:376-402)"""
    if field_values.ndim == 2:
        v = np.asarray(value, dtype=np.float64)
        if v.shape != (3,):
            raise ValueError("Vectors must have 3 components")
        field_values[:] = v[:, None]
    else:
        field_values[:] = float(value)


# -------------------------------------------------- boundary conditions
@dataclass
class _BC:
    ID: str
    value: Any = 0.0
    kind: int = -1


def Dirichlet(name, value): return _BC(name, value, A.BC_DIRICHLET)
def Neumann(name, value=0.0): return _BC(name, value, A.BC_NEUMANN)
def Zerogradient(name): return _BC(name, 0.0, A.BC_ZEROGRADIENT)
def Extrapolated(name): return _BC(name, 0.0, A.BC_EXTRAPOLATED)
def Wall(name, value=0.0): return _BC(name, value, A.BC_WALL)
def Symmetry(name): return _BC(name, 0.0, A.BC_SYMMETRY)
def Empty(name): return _BC(name, 0.0, A.BC_EMPTY)
# Outlet (outlet.jl): zero gradient in the pressure-based solvers
def Outlet(name): return _BC(name, 0.0, A.BC_ZEROGRADIENT)
# wall functions with the reference's default coefficients (Discretise_0_types.jl:63,108,154)
def KWallFunction(name): return _BC(name, 0.0, A.BC_KWALL)
def OmegaWallFunction(name): return _BC(name, 0.0, A.BC_OMEGAWALL)
def NutWallFunction(name): return _BC(name, 0.0, A.BC_NUTWALL)


def assign(region, **fields):
    """assign(region, (U=[...], p=[...])): resolve patch names, require one BC
    per patch (Discretise_4_assign_boundaries.jl:5-49)."""
    out = {}
    for name, bcs in fields.items():
        resolved = []
        for bc in bcs:
            patch = region.patch_index(bc.ID)  # KeyError: not a recognised boundary name
            v = bc.value
            if np.ndim(v):
                if len(v) != 3:
                    raise ValueError("Vectors must have 3 components")
                v = [float(x) for x in v]
            else:
                v = float(v)
            resolved.append((bc.kind, patch, v))
        if len(resolved) != len(region.patch_names):
            raise AssertionError(
                f"Inconsistent number of boundaries assigned to field {name}")
        out[name] = resolved
    return out


def simple_setup(model: Physics, config: Configuration, pref=None,
                 ncorrectors: int = 0) -> A.SimpleSetupC:
    """xcal_simple_setup for run!(model, config; pref, ncorrectors)."""
    if not isinstance(model.turbulence, (Laminar, KOmega)):
        raise NotImplementedError("laminar and k-omega SIMPLE are on the device path")
    sU, sp = config.schemes["U"], config.schemes["p"]
    if sU.time not in TIME:
        raise NotImplementedError(f"time scheme {sU.time!r}")
    for s in (sU, sp):
        if s.gradient not in _GRAD or s.limiter is not None or s.laplacian is not Linear:
            raise NotImplementedError("device path: Gauss or Midpoint gradient, Linear "
                                      "laplacian, no limiter")
    div = _DIV.get(sU.divergence)
    if div is None:
        raise NotImplementedError(f"divergence scheme {sU.divergence!r}")
    bU = A.bc_array(config.boundaries["U"])
    bp = A.bc_array(config.boundaries["p"])
    setup = A.SimpleSetupC(float(model.fluid.nu), div, int(pref is not None),
                           float(pref or 0.0), config.solvers["U"].c(), config.solvers["p"].c(),
                           bU, bp, len(config.boundaries["U"]), _GRAD[sp.gradient],
                           _GRAD[sU.gradient], TIME[sU.time], float(config.runtime.time_step),
                           int(ncorrectors))
    setup._keep = (bU, bp)
    return setup


def turbulence_setup(model: Physics, config: Configuration):
    """xcal_turbulence_setup of RANS{KOmega} (None for laminar)."""
    t = model.turbulence
    if not isinstance(t, KOmega):
        return None
    sk = config.schemes.get("k", Schemes())
    div = _DIV.get(sk.divergence)
    if div is None:
        raise NotImplementedError(f"divergence scheme {sk.divergence!r}")
    bk = A.bc_array(config.boundaries["k"])
    bw = A.bc_array(config.boundaries["omega"])
    bn = A.bc_array(config.boundaries["nut"])
    ts = A.TurbulenceSetupC(A.TURB_KOMEGA, t.beta_star, t.alpha1, t.beta1, t.sigma_k,
                            t.sigma_omega, div, config.solvers["k"].c(),
                            config.solvers["omega"].c(), bk, bw, bn)
    ts._keep = (bk, bw, bn)
    return ts


_ctx_cache = {}


def _context(device: int) -> hip.Context:
    if device not in _ctx_cache:
        _ctx_cache[device] = hip.Context(device)
    return _ctx_cache[device]


def save_output(model: Physics, outputWriter, iteration: int, time, config: Configuration,
                directory=None):
    """save_output(model, outputWriter, iteration, time, config): U and p,
    plus k, omega and nut for k-omega (RANS_laminar.jl:91-110,
    RANS_kOmega.jl:235-260).  Returns the path written."""
    args = [("U", model.momentum.U), ("p", model.momentum.p)]
    if isinstance(model.turbulence, KOmega):
        t = model.turbulence
        args += [("k", t.k), ("omega", t.omega), ("nut", t.nut)]
    return write_results(iteration, time, model.domain, outputWriter, config.boundaries, *args,
                         directory=directory)


def run(model: Physics, config: Configuration, pref=None, ncorrectors: int = 0,
        inner_loops: Optional[int] = None, output=VTK(), directory=None):
    """run!(model, config; output, pref, ncorrectors, inner_loops) for
    incompressible flow: SIMPLE for a Steady model (laminar or k-omega), PISO
    with `inner_loops` pressure correctors per time step (default 2) for a
    Transient one (laminar), as run! calls piso!(model, config; pref,
    inner_loops) (Solvers_3_solver_dispatch.jl:186-190, 224-235).  Returns
    the residual histories (Ux, Uy, Uz, p) and leaves the final fields in
    model.momentum.  The SIMPLE loop stops where the device's convergence
    test (every residual, a turbulence model's included) stops it.

    Results are written as the reference writes them: with
    config.runtime.write_interval w >= 0, after every iteration that is a
    multiple of w, and (SIMPLE) at the iteration that converges
    (Solvers_1_SIMPLE.jl:234-258, Solvers_2_PISO.jl:205-208); w < 0 writes
    nothing.  The device loop runs the iterations between two writes in one
    call and stays resident across calls, so the iterates are those of one
    uninterrupted run.  `directory` (default: the working directory) is
    where the files go; with output=OpenFOAM() it is the case directory,
    whose constant/polyMesh is written at the start of the run."""
    if ncorrectors < 0:
        raise ValueError("ncorrectors must be >= 0")
    if inner_loops is None:
        inner_loops = 2
    elif inner_loops < 1:
        raise ValueError("inner_loops must be >= 1")
    if not isinstance(config.hardware.backend, HIPBackend):
        raise TypeError("the device path runs on HIPBackend()")
    wi = config.runtime.write_interval
    if wi == 0:
        raise ZeroDivisionError("write_interval 0: iteration % write_interval is undefined")
    ctx = _context(config.hardware.backend.device)
    dmesh = hip.DeviceMesh(ctx, model.domain)
    sim = hip.Simple(dmesh, simple_setup(model, config, pref, ncorrectors))
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    ts = turbulence_setup(model, config)
    if ts is not None:
        sim.set_turbulence(ts)
        t = model.turbulence
        sim.set_scalar("k", t.k)
        sim.set_scalar("omega", t.omega)
        sim.set_scalar("nut", t.nut)
    transient = isinstance(model.time, Transient)
    if transient and ts is not None:
        raise NotImplementedError("transient k-omega is not on the device path")
    if isinstance(output, OpenFOAM):  # writes constant/polyMesh now, as run! does
        writer = initialise_writer(output, model.domain, directory)
    else:
        writer = initialise_writer(output, model.domain) if wi > 0 else None
    if wi < 0:
        writer = None

    def fetch():
        model.momentum.U[:] = sim.get("U")
        model.momentum.p[:] = sim.get("p")
        if ts is not None:
            for name in ("k", "omega", "nut"):
                getattr(model.turbulence, name)[:] = sim.get(name)

    iterations = config.runtime.iterations
    step = iterations if writer is None else wi
    parts, it, time = [], 0, 0.0
    while it < iterations:
        chunk = min(step - it % step, iterations - it)
        out = sim.run_piso(chunk, inner_loops) if transient else sim.run(chunk)
        parts.append(out["residuals"])
        done = out["done"]
        if transient:
            for _ in range(done):
                time += config.runtime.time_step
        it += done
        stop = not transient and done > 0 and out["converged"]
        if writer is not None and (stop or it % wi == 0):
            fetch()
            save_output(model, writer, it, time if transient else it, config, directory)
        if stop or done < chunk:
            break
    fetch()
    r = np.concatenate(parts, axis=1) if parts else np.zeros((4, 0))
    return dict(Ux=r[0], Uy=r[1], Uz=r[2], p=r[3])
