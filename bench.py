#!/usr/bin/env python3
"""SIMPLE cell-iterations per second on one MI355X.

Workload (a "step" is one SIMPLE iteration over the whole mesh):
  3D laminar channel, steady incompressible SIMPLE, FP64, on a
  100x100x100 hex box (1M cells) whose interior nodes are jittered by +-10 %
  of the spacing and whose cells are renumbered by a seeded random
  permutation.  Inlet U = (1,0,0), outlet p = 0, no-slip walls, nu = 1e-3,
  upwind convection, linear Laplacian, Gauss gradients.  U: BiCGSTAB + DILU
  (rtol 0.1), relax 0.7; p: CG + DILU to rtol 1e-10, relax 0.3.

value     = cells x steps / seconds of the timed region (inputs resident in HBM)
roofline  = algorithmic bytes of the dominant pressure-solve kernel per launch
            / its mean launch time from HIP events on the library's stream;
            the dominant one of the forward DILU sweep, the backward sweep and
            the CG SpMV (most time per CG iteration), the other two under
            roofline.others
cpu_baseline = the CPU oracle (restatement of the reference's FP64 path)
            timed on the host's cores over a bounded sample of the same run.
detail.mixed_precision = the same channel with FP32 inner Krylov solves in
            FP64 iterative refinement, and its per-field error against FP64;
detail.batched = 256 batched 2D cavities (FP64 and FP32-IR pressure), the
            form that --mode batch reports in full: value with eager
            launches (the default), the same run with the Krylov iterations
            replayed from HIP graphs beside it, and the rows per thread the
            built-in tuner picked for the DILU sweeps with its timings.
detail.serving = 8 k-omega step cases served with Poisson arrivals and
            continuous batching: p50/p99 time to converged over the converged
            requests, with the counts of converged, capped and diverged ones;
            --mode serve reports it in full.  --quick skips these three.
detail.config1 = the reference's CPU configuration (40x40 cavity, 200
            iterations): the CPU oracle (median of 3) beside the device.

Other modes: --mode batch (batched cavities), --mode serve (k-omega serving
with Poisson arrivals: p50/p99 time to converged), --mode stepbatch (the
serving workload's first --cases requests stepped as one batch), --mode piso (the channel as
a transient PISO case: cell-steps/s), --mode calib (a plain 256 MiB copy for
HBM counter calibration).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "SIMPLE cell-iters/s FP64 vs mixed@1e-8; batched solves/s + p50/p99 latency"
SEED_JITTER, SEED_PERM = 0x5EED0002, 0x5EED0003
SEED_BATCH = 0x5EED0004
SEED_SERVE = 0x5EED0005


def step_case(api, x, ratio, cells, iterations, U_in=10.0, nu=None, rans=True, Tu=0.05,
              nut_ratio=100.0, precon="DILU"):
    """2D backward-facing step from the block mesher: step height `ratio` of
    the outlet height 1, inlet length 1, outlet length 4, about `cells` cells.
    Inlet U = (U_in, 0, 0), outlet p = 0, walls.  rans: k-omega with wall
    functions, inlet k = 1.5 (Tu U_in)^2 and omega = k / (nut_ratio nu), every
    field started from its inlet value (nu 1e-5 by default); otherwise laminar
    from rest (nu 0.05 by default).  All solves BiCGSTAB (CG for p) with the
    `precon` preconditioner, upwind convection, residual target 1e-6."""
    nu = nu if nu is not None else (1e-5 if rans else 0.05)
    ny = max(4, int(round(np.sqrt(cells / (5.0 - ratio)))))
    ny_low = max(1, int(round(ratio * ny)))
    ny_up = max(1, ny - ny_low)
    mesh = x.mesh.backward_facing_step_2d(ratio, 1.0, 1.0, 4.0, nx_in=ny, nx_out=4 * ny,
                                          ny_low=ny_low, ny_up=ny_up)
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=nu),
                        turbulence=api.RANS(api.KOmega) if rans else api.RANS(api.Laminar))
    bc = dict(U=[api.Dirichlet("inlet", [U_in, 0.0, 0.0]), api.Extrapolated("outlet"),
                 api.Wall("walls", [0, 0, 0]), api.Wall("top", [0, 0, 0])])

    def S(solver, relax, rtol=None):
        tol = {} if rtol is None else dict(rtol=rtol, atol=1e-10)
        return api.SolverSetup(solver=solver, preconditioner=getattr(api, precon)(), convergence=1e-6,
                               relax=relax, **tol)
    if rans:
        k_in = 1.5 * (Tu * U_in) ** 2
        w_in = k_in / (nut_ratio * nu)
        for f, v in ((model.momentum.U, [U_in, 0.0, 0.0]), (model.turbulence.k, k_in),
                     (model.turbulence.omega, w_in), (model.turbulence.nut, k_in / w_in)):
            api.initialise(f, v)
        bc.update(
            p=[api.Neumann("inlet", 0.0), api.Dirichlet("outlet", 0.0), api.Wall("walls"),
               api.Wall("top")],
            k=[api.Dirichlet("inlet", k_in), api.Extrapolated("outlet"),
               api.KWallFunction("walls"), api.KWallFunction("top")],
            omega=[api.Dirichlet("inlet", w_in), api.Extrapolated("outlet"),
                   api.OmegaWallFunction("walls"), api.OmegaWallFunction("top")],
            nut=[api.Dirichlet("inlet", k_in / w_in), api.Extrapolated("outlet"),
                 api.NutWallFunction("walls"), api.NutWallFunction("top")])
        solvers = dict(U=S(api.Bicgstab(), 0.7, 1e-2), p=S(api.Cg(), 0.3, 1e-3),
                       k=S(api.Bicgstab(), 0.7, 1e-2), omega=S(api.Bicgstab(), 0.7, 1e-2))
    else:
        bc.update(p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                     api.Extrapolated("walls"), api.Extrapolated("top")])
        solvers = dict(U=S(api.Bicgstab(), 0.7), p=S(api.Cg(), 0.3))
    config = api.Configuration(
        solvers=solvers, schemes={f: api.Schemes(divergence=api.Upwind) for f in solvers},
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=api.assign(region=mesh, **bc))
    return mesh, model, config


def submit_case(srv, api, mesh, model, config):
    """Submit one step case to a Server (turbulent cases carry k, omega, nut)."""
    setup = api.simple_setup(model, config)
    t = model.turbulence
    if getattr(t, "k", None) is None:
        return srv.submit(mesh, setup, model.momentum.U, model.momentum.p)
    return srv.submit(mesh, setup, model.momentum.U, model.momentum.p,
                      turbulence=api.turbulence_setup(model, config), k0=t.k, omega0=t.omega,
                      nut0=t.nut)


def serve_cases(nreq, seed=SEED_SERVE):
    """Step-height ratio ~ U[0.3, 0.6], cells log-uniform in [4k, 250k]."""
    rng = np.random.default_rng(seed)
    ratio = rng.uniform(0.3, 0.6, nreq)
    cells = np.exp(rng.uniform(np.log(4e3), np.log(2.5e5), nreq))
    return ratio, cells


def serve_leg(args, x, requests, rate, max_iterations, alone=True):
    """Serving form: `requests` step cases arrive as a Poisson process at
    `rate` per second; the server admits them by HBM footprint into its
    running batch (continuous batching: the cases in flight step together)
    and each runs to residual 1e-6 or `max_iterations`.  Returns
    (cell-iterations per second over the makespan, summary dict)."""
    api, H = x.api, x.hip
    ratio, cells = serve_cases(requests)
    t0 = time.perf_counter()
    cases = [step_case(api, x, r, c, max_iterations, rans=not args.laminar, precon=args.precon)
             for r, c in zip(ratio, cells)]
    t_mesh = time.perf_counter() - t0
    rng = np.random.default_rng(SEED_SERVE + 1)
    gaps = rng.exponential(1.0 / rate, requests)
    budget = int(args.hbm_gb * (1 << 30))
    srv = H.Server(0, slots=args.slots, hbm_budget=budget, max_iterations=max_iterations,
                   lanes=args.lanes)
    ids, submit_s = [], []
    t0 = time.perf_counter()
    for (mesh, model, config), gap in zip(cases, gaps):
        if ids:
            time.sleep(gap)  # exponential inter-arrival gap
        ts = time.perf_counter()
        ids.append(submit_case(srv, api, mesh, model, config))
        submit_s.append(time.perf_counter() - ts)
        log(f"submitted {len(ids)}/{requests} at {time.perf_counter() - t0:.2f} s")
    while not srv.wait(0, 30.0):
        done = sum(srv.status(i)["state"] >= x._abi.REQ_DONE for i in ids)
        log(f"{done}/{len(ids)} finished at {time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    st = [srv.status(i) for i in ids]
    lat = np.array([s["t_end"] - s["t_submit"] for s in st])
    wait = np.array([s["t_start"] - s["t_submit"] for s in st])
    its = np.array([s["iterations"] for s in st])
    ncell = np.array([s["cells"] for s in st])
    ok = np.array([s["state"] == x._abi.REQ_DONE for s in st])
    conv = np.array([s["converged"] for s in st]) & ok
    mix = {i: srv.fetch(i) for i, good in zip(ids, ok) if good}
    rebuilds = max(s["batch_rebuilds"] for s in st)
    srv.close()
    # each request again, alone on a one-slot server: its run time without
    # the mix, and whether its fields are the same as in the mix
    alone_s, same = [], True
    if alone:
        srv1 = H.Server(0, slots=1, hbm_budget=budget, max_iterations=max_iterations)
        for k, ((mesh, model, config), rid) in enumerate(zip(cases, ids)):
            j = submit_case(srv1, api, mesh, model, config)
            srv1.wait(j, 0.0)
            s1 = srv1.status(j)
            alone_s.append(s1["t_end"] - s1["t_start"])
            same = same and s1["iterations"] == st[k]["iterations"] and s1["state"] == st[k]["state"]
            if rid in mix:
                got = srv1.fetch(j)
                same = same and all(np.array_equal(got[f], mix[rid][f]) for f in got)
            srv1.release(j)
            log(f"alone {k + 1}/{len(ids)}: {alone_s[-1]:.2f} s")
        srv1.close()
    pct = lambda a, q: round(float(np.percentile(a, q)), 3) if len(a) else None
    big = int(np.argmax(ncell))
    # admission cost of the largest request's matrix connectivity (CSR
    # pattern, slot maps, DILU levels): device kernels against the host loops
    ctx = H.Context(0)
    Kd = [H.Connectivity(ctx, cases[big][0], on_device=True) for _ in range(2)][-1]
    Kh = H.Connectivity(ctx, cases[big][0], on_device=False)
    ingest = {"cells": int(ncell[big]), "levels": list(Kd.levels),
              "device_build_ms": round(Kd.wall_ms, 2), "device_kernels_ms": round(Kd.device_ms, 2),
              "host_build_ms": round(Kh.wall_ms, 2)}
    del Kd, Kh, ctx
    summary = {
        "workload": f"{requests} 2D backward-facing steps, step ratio U[0.3,0.6], cells "
                    f"log-uniform [4k,250k], "
                    + ("laminar (nu 0.05" if args.laminar else "k-omega RANS (nu 1e-5")
                    + f", U 10), {args.precon}, Poisson arrivals at {rate}/s, residual 1e-6, "
                    f"at most {max_iterations} iterations",
        "slots": args.slots, "lanes": args.lanes, "hbm_budget_gb": args.hbm_gb,
        "converged": int(conv.sum()), "capped": int(np.sum(ok & ~conv)),
        "diverged": int(np.sum(~ok)),
        "p50_time_to_converged_s": pct(lat[conv], 50),
        "p99_time_to_converged_s": pct(lat[conv], 99),
        "p50_time_to_stop_all_s": pct(lat, 50), "p99_time_to_stop_all_s": pct(lat, 99),
        "queue_wait_p50_s": pct(wait, 50), "queue_wait_max_s": pct(wait, 100),
        "submit_call_p50_s": pct(np.array(submit_s), 50),
        "submit_call_largest_case_s": round(submit_s[big], 3),
        "connectivity_largest_case": ingest,
        "largest_case_cells": int(ncell[big]),
        "makespan_s": round(wall, 2), "requests_per_s": round(requests / wall, 3),
        "batch_rebuilds": int(rebuilds), "mesh_build_s": round(t_mesh, 2),
        "alone_p50_s": pct(alone_s, 50) if alone_s else None,
        "alone_p99_s": pct(alone_s, 99) if alone_s else None,
        "identical_alone_vs_mix": same if alone_s else None,
        "iterations": its.tolist(), "cells": ncell.tolist(),
        "converged_flags": conv.astype(int).tolist(),
        "lane_of_request": [s["lane"] for s in st],
    }
    return float(np.sum(ncell * its)) / wall, summary


def run_stepbatch(args, x):
    """--mode stepbatch: the serving workload's first `--cases` requests
    stepped together as one batch mesh (no arrivals, no server): k-omega
    SIMPLE iterations per second of the batch the server would form, with
    per-stage device time and Krylov counts."""
    api, H = x.api, x.hip
    k = min(args.cases, 8)
    ratio, cells = serve_cases(max(8, k))
    cases = [step_case(api, x, r, c, args.warmup + args.steps, rans=not args.laminar,
                       precon=args.precon) for r, c in zip(ratio[:k], cells[:k])]
    meshes = [c[0] for c in cases]
    ctx = H.Context(0)
    dm = H.DeviceMesh.batch(ctx, meshes)
    model, config = cases[0][1], cases[0][2]
    sim = H.Simple(dm, api.simple_setup(model, config))
    if not args.laminar:
        sim.set_turbulence(api.turbulence_setup(model, config))
        for f in ("k", "omega", "nut"):
            sim.set_scalar(f, np.concatenate([np.asarray(getattr(c[1].turbulence, f),
                                                         np.float64).reshape(-1) for c in cases]))
    sim.set_U(np.concatenate([np.asarray(c[1].momentum.U, np.float64).reshape(3, -1)
                              for c in cases], axis=1))
    if args.warmup:
        sim.run_batch(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = sim.run_batch(args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kit = out["krylov_iters"]
    line = {
        "metric": METRIC, "value": round(dm.n * out["done"] / wall, 1), "unit": "cell-iters/s",
        "n_gpus": 1, "steps": out["done"], "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / max(out["done"], 1), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (the serving workload's seeded step ratios and sizes)",
        "config": {"workload": f"first {k} serving requests as one batch: 2D backward-facing "
                               f"steps, " + ("laminar" if args.laminar else "k-omega RANS")
                               + f", {args.precon}",
                   "cells": dm.n, "case_cells": [m.n_cells for m in meshes],
                   "parallelism": "single GPU, one batch mesh"},
        "roofline": None, "cpu_baseline": None,
        "detail": {"device_ms": {kk: round(v, 2) for kk, v in sim.stage_times().items()},
                   "krylov_iters_mean_per_case_Ux_Uy_Uz_p": kit.mean(axis=2).round(1).tolist()},
    }
    print(json.dumps(line))


def run_serve(args, x):
    """--mode serve: the serving leg as the bench line."""
    value, summary = serve_leg(args, x, args.requests, args.rate, args.max_iterations,
                               alone=not args.no_alone)
    line = {
        "metric": METRIC, "value": round(value, 1),
        "unit": "cell-iters/s", "n_gpus": 1, "steps": int(sum(summary["iterations"])), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded step ratios, sizes and arrival times)",
        "config": {"workload": "serving " + summary["workload"],
                   "slots": args.slots, "hbm_budget_gb": args.hbm_gb, "precon": args.precon,
                   "max_iterations": args.max_iterations, "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
        "detail": {"serving": summary},
    }
    print(json.dumps(line))


def cavity(api, mesh, iterations, precision="fp64"):
    """Lid-driven cavity settings of the reference's 2D laminar example: lid
    U = (1,0,0), no-slip walls, p Wall everywhere with pref = 0 in cell 1,
    upwind, U: BiCGSTAB+DILU relax 0.7, p: CG+DILU relax 0.3, rtol 0.1."""
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=0.01))
    BCs = api.assign(region=mesh,
                     U=[api.Wall("bottom", [0, 0, 0]), api.Wall("top", [1, 0, 0]),
                        api.Wall("left", [0, 0, 0]), api.Wall("right", [0, 0, 0])],
                     p=[api.Wall("bottom"), api.Wall("top"), api.Wall("left"), api.Wall("right")])
    solvers = dict(U=api.SolverSetup(solver=api.Bicgstab(), preconditioner=api.DILU(),
                                     convergence=1e-7, relax=0.7),
                   p=api.SolverSetup(solver=api.Cg(), preconditioner=api.DILU(),
                                     convergence=1e-7, relax=0.3, precision=precision))
    config = api.Configuration(
        solvers=solvers, schemes=dict(U=api.Schemes(divergence=api.Upwind), p=api.Schemes()),
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config


def config1(x, O, threads, runs=3):
    """The reference's CPU configuration: the 40x40 lid-driven cavity (Re 100,
    pref 0 in cell 1), 200 SIMPLE iterations, FP64.  The CPU oracle (the
    restatement of the reference's FP64 path) on `threads` OpenMP threads,
    median of `runs` runs after one warm-up, beside the device running the
    same case the same way."""
    api, H = x.api, x.hip
    mesh = x.mesh.cavity2d(40)
    iters = 200
    O.set_threads(threads)
    om = O.OracleMesh(mesh, x._abi)
    times = []
    for r in range(runs + 1):
        model, config = cavity(api, mesh, iters)
        setup = api.simple_setup(model, config, pref=0.0)
        t0 = time.perf_counter()
        ref = om.simple_run(setup, model.momentum.U, model.momentum.p, iters)
        if r > 0:
            times.append(time.perf_counter() - t0)
    tc = float(np.median(times))
    ctx = H.Context(0)
    dm = H.DeviceMesh(ctx, mesh)
    gtimes = []
    for r in range(runs + 1):
        model, config = cavity(api, mesh, iters)
        sim = H.Simple(dm, api.simple_setup(model, config, pref=0.0))
        sim.set_U(model.momentum.U)
        sim.set_p(model.momentum.p)
        ctx.synchronize()
        t0 = time.perf_counter()
        out = sim.run(iters)
        ctx.synchronize()
        if r > 0:
            gtimes.append(time.perf_counter() - t0)
        del sim
    tg = float(np.median(gtimes))
    n = mesh.n_cells
    return {"workload": "2D lid-driven cavity, 40x40 quads (1600 cells), Re 100, 200 SIMPLE "
                        "iterations, FP64",
            "cpu": {"value": round(n * ref["done"] / tc, 1), "unit": "cell-iters/s",
                    "cores": threads, "kind": "port", "median_s": round(tc, 4), "runs": runs},
            "gpu": {"value": round(n * out["done"] / tg, 1), "unit": "cell-iters/s",
                    "median_s": round(tg, 4), "runs": runs}}


def batch_cases(ncase, seed=SEED_BATCH):
    """Cavity sizes m = round(32 * 8^u), u ~ U[0,1), and Re ~ U[50, 400]."""
    rng = np.random.default_rng(seed)
    m = np.round(32.0 * 8.0 ** rng.random(ncase)).astype(int)
    Re = rng.uniform(50.0, 400.0, ncase)
    return m, Re


def run_batch(args, x):
    print(json.dumps(batch_line(args, x, args.precision)))


def batch_line(args, x, precision, steps=None, warmup=None):
    """Batched form: independent 2D cavities stepped together on one batch mesh."""
    api, H = x.api, x.hip
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    sizes, Re = batch_cases(args.cases)
    t0 = time.perf_counter()
    meshes = [x.mesh.cavity2d(int(m)) for m in sizes]
    t_mesh = time.perf_counter() - t0
    model, config = cavity(api, meshes[0], warmup + steps, precision)
    setup = api.simple_setup(model, config, pref=0.0)
    ctx = H.Context(0)
    t0 = time.perf_counter()
    dm = H.DeviceMesh.batch(ctx, meshes)
    ctx.synchronize()
    t_upload = time.perf_counter() - t0

    def timed_run(graph, kernel_timing):
        """warmup + steps on a fresh SIMPLE state; seconds of the timed steps.
        graph: Krylov iterations replayed from HIP graphs (otherwise eager)."""
        os.environ["XCAL_GRAPH"] = "1" if graph else "0"
        try:
            sm = H.Simple(dm, setup)
            sm.set_case_nu(1.0 / Re)
            if warmup:
                sm.run_batch(warmup)
            if kernel_timing:
                ctx.time_kernel(args.roofline_kernel, True)
            ctx.synchronize()
            ta = time.perf_counter()
            o = sm.run_batch(steps)
            ctx.synchronize()
            tb = time.perf_counter()
        finally:
            os.environ.pop("XCAL_GRAPH", None)
        if o["done"] != steps:
            raise RuntimeError(f"batch stopped after {o['done']} of {steps} steps")
        return sm, o, tb - ta

    # the same run from the same start twice: HIP-graph replay of the Krylov
    # iterations (XCAL_GRAPH=1), then eager launches (the library's default,
    # the headline value) with the roofline kernel timed
    _, _, wall_graph = timed_run(True, False)
    sim, out, wall = timed_run(False, True)
    launches, kms = ctx.kernel_time(args.roofline_kernel)
    ctx.time_kernel(args.roofline_kernel, False)
    n = dm.n
    f_rows, b_rows, tune_us = dm.pipe_tuning()
    vb = 4 if precision == "fp32-ir" else 8
    nbytes = kernel_bytes(args.roofline_kernel, n, dm.nnz, vb)
    avg_s = kms / 1e3 / max(launches, 1)
    achieved = nbytes / avg_s / 1e9
    err = None
    if precision != "fp64":
        m64, c64 = cavity(api, meshes[0], warmup + steps)
        s64 = H.Simple(dm, api.simple_setup(m64, c64, pref=0.0))
        s64.set_case_nu(1.0 / Re)
        s64.run_batch(warmup + steps)

        def rel(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        err = {f: rel(sim.get(f), s64.get(f)) for f in ("U", "p", "mdotf")}
    kit = out["krylov_iters"]
    line = {
        "metric": METRIC, "value": round(n * steps / wall, 1), "unit": "cell-iters/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * wall / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f64 (pressure Krylov f32 in f64 refinement)",
        "data": "synthetic (seeded cavity sizes and Reynolds numbers)",
        "config": {"workload": f"{args.cases} batched 2D lid-driven cavities, m x m quads with m "
                               f"log-uniform in [32,256], Re uniform in [50,400], SIMPLE "
                               f"{precision.upper()} pressure",
                   "cases": args.cases, "cells": n, "nnz": dm.nnz,
                   "p_solver": "CG+DILU rtol 0.1", "U_solver": "BiCGSTAB+DILU rtol 0.1",
                   "parallelism": "single GPU, one batch mesh"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": hbm_traffic(args.roofline_kernel,
                                            f"batch{args.cases}_{precision}"),
                     "kernel": args.roofline_kernel, "bytes_per_launch": nbytes,
                     "timed_launches": launches, "avg_us": round(avg_s * 1e6, 2)},
        "cpu_baseline": None,
        "detail": {"mesh_build_s": round(t_mesh, 2), "upload_s": round(t_upload, 2),
                   "case_iters_per_s": round(args.cases * steps / wall, 1),
                   "graph": {"on_ms_per_step": round(1e3 * wall_graph / steps, 3),
                             "off_ms_per_step": round(1e3 * wall / steps, 3),
                             "speedup": round(wall / wall_graph, 4), "default": "off"},
                   "tuned": {"dilu_rows_per_thread": [f_rows, b_rows],
                             "tuner_us_per_launch": {f"{d}_rw{r}": round(v, 2)
                                                     for (d, r), v in tune_us.items()}},
                   "device_ms": {k: round(v, 2) for k, v in sim.stage_times().items()},
                   "p_iters_max_mean": float(np.mean(kit[:, 3].max(axis=0))),
                   "error_vs_fp64": err},
    }
    return line


def channel(api, mesh, iterations, precision="fp64", inner_rtol=1e-6):
    model = api.Physics(domain=mesh, fluid=api.Fluid(nu=1e-3))
    api.initialise(model.momentum.U, [1.0, 0.0, 0.0])
    api.initialise(model.momentum.p, 0.0)
    BCs = api.assign(region=mesh,
                     U=[api.Dirichlet("inlet", [1.0, 0.0, 0.0]), api.Extrapolated("outlet"),
                        api.Wall("walls", [0.0, 0.0, 0.0])],
                     p=[api.Extrapolated("inlet"), api.Dirichlet("outlet", 0.0),
                        api.Extrapolated("walls")])
    solvers = dict(U=api.SolverSetup(solver=api.Bicgstab(), preconditioner=api.DILU(),
                                     convergence=1e-7, relax=0.7, rtol=0.1),
                   p=api.SolverSetup(solver=api.Cg(), preconditioner=api.DILU(),
                                     convergence=1e-7, relax=0.3, rtol=1e-10,
                                     precision=precision, inner_rtol=inner_rtol))
    config = api.Configuration(
        solvers=solvers, schemes=dict(U=api.Schemes(divergence=api.Upwind), p=api.Schemes()),
        runtime=api.Runtime(iterations=iterations, write_interval=-1),
        hardware=api.Hardware(backend=api.HIPBackend(0)), boundaries=BCs)
    return model, config


def transient_channel(api, mesh, steps, dt):
    """The channel above as a transient PISO case: Time{Euler} with step dt,
    U BiCGSTAB + DILU (rtol 1e-3, unrelaxed), p CG + DILU (rtol 1e-6), two
    pressure correctors per step."""
    model, config = channel(api, mesh, steps)
    model.time = api.Transient()
    config.schemes["U"] = api.Schemes(divergence=api.Upwind, time=api.Euler)
    config.runtime.time_step = dt
    config.solvers["U"] = api.SolverSetup(solver=api.Bicgstab(), preconditioner=api.DILU(),
                                          convergence=1e-7, relax=1.0, rtol=1e-3)
    config.solvers["p"] = api.SolverSetup(solver=api.Cg(), preconditioner=api.DILU(),
                                          convergence=1e-7, relax=1.0, rtol=1e-6)
    return model, config


def run_piso(args, x):
    """Transient PISO on the 1M-cell channel: cell-steps per second, with the
    CG SpMV roofline as in the channel line."""
    api, H = x.api, x.hip
    N = args.cells
    t0 = time.perf_counter()
    mesh = x.mesh.box3d(N, N, N, jitter=0.1, seed_jitter=SEED_JITTER, seed_perm=SEED_PERM)
    t_mesh = time.perf_counter() - t0
    dt = 0.005
    model, config = transient_channel(api, mesh, args.warmup + args.steps, dt)
    setup = api.simple_setup(model, config)
    ctx = H.Context(0)
    t0 = time.perf_counter()
    dm = H.DeviceMesh(ctx, mesh)
    sim = H.Simple(dm, setup)
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    ctx.synchronize()
    t_upload = time.perf_counter() - t0
    if args.warmup:
        sim.run_piso(args.warmup, 2)
    ctx.time_kernel(args.roofline_kernel, True)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = sim.run_piso(args.steps, 2)
    ctx.synchronize()
    t1 = time.perf_counter()
    launches, kms = ctx.kernel_time(args.roofline_kernel)
    ctx.time_kernel(args.roofline_kernel, False)
    wall = t1 - t0
    n = mesh.n_cells
    nbytes = kernel_bytes(args.roofline_kernel, n, dm.nnz)
    avg_s = kms / 1e3 / max(launches, 1)
    achieved = nbytes / avg_s / 1e9
    kit = out["krylov_iters"]
    line = {
        "metric": "PISO cell-steps/s (transient channel, 2 pressure correctors per step)",
        "value": round(n * args.steps / wall, 1), "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered, renumbered hex box)",
        "config": {"workload": f"3D laminar channel PISO FP64, {N}x{N}x{N} jittered+renumbered "
                               f"hex box ({n} cells), Euler dt {dt}",
                   "cells": n, "nnz": dm.nnz, "inner_loops": 2,
                   "p_solver": "CG+DILU rtol 1e-6", "U_solver": "BiCGSTAB+DILU rtol 1e-3",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": hbm_traffic(args.roofline_kernel, f"channel_{N}cubed_fp64"),
                     "kernel": args.roofline_kernel, "bytes_per_launch": nbytes,
                     "timed_launches": launches, "avg_us": round(avg_s * 1e6, 2)},
        "cpu_baseline": None,
        "detail": {"mesh_build_s": round(t_mesh, 2), "upload_s": round(t_upload, 2),
                   "p_cg_iters_mean_last_corrector": float(np.mean(kit[3])),
                   "U_bicgstab_iters_mean": float(np.mean(kit[0]))},
    }
    print(json.dumps(line))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_calib(args):
    """HBM counter calibration: a plain 256 MiB device-to-device copy, three
    times.  rocprofv3 --pmc of the TCC_EA0_RDREQ_* counters and of WRITE_SIZE
    on this command give the counter readings of a kernel that moves a known
    number of bytes."""
    import torch
    nbytes = 256 << 20
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "launches": 3}))


def hbm_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (L2-to-memory read requests by size, plus
    WRITE_SIZE, calibrated against a copy kernel), or None when no summary
    covers this pair."""
    import re
    path = ROOT / "profiles" / f"hbm_traffic_{kernel}_{workload}.md"
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def kernel_bytes(name, n, nnz, vb=8, vv=None):
    """Algorithmic HBM bytes of one timed unit (vb-byte values, int32
    indices; every array element counted once, gathers included once, slice
    padding and the cone form's recomputed rows not counted).  cg_spmv is
    one launch, counted as a CSR SpMV whatever the storage: nnz (value +
    index) + (n + 1) row pointers + x once + y.  dilu_fwd and dilu_bwd are
    one DILU sweep each (one launch in the cone form), counted as a CSR
    triangular sweep: the strict triangle's nlow = (nnz - n) / 2 entries
    (value + index), n + 1 row pointers, n diagonal pointers, and rhs, D
    and out (3 n values); plus the work fused into them: the forward
    sweep's CG update (x, p, q read; x, r written: 5 n values) and the
    backward sweep's r.z (r read: n values).  vv is the Krylov vectors'
    value size where it differs from the factors' (FP32 sweeps inside FP64
    CG): the forward sweep's rhs r and the backward sweep's z and r are
    vectors; y and D are the sweeps' own."""
    vv = vb if vv is None else vv
    nlow = (nnz - n) // 2
    if name == "cg_spmv":
        return (4 + vb) * nnz + 4 * (n + 1) + vb * 2 * n
    sweep = (vb + 4) * nlow + 4 * (n + 1) + 4 * n + 2 * vb * n + vv * n
    if name == "dilu_fwd":
        return sweep + 5 * vv * n
    if name == "dilu_bwd":
        return sweep + vv * n
    raise KeyError(name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=100, help="cells per box edge")
    ap.add_argument("--roofline-kernel", default="cg_spmv",
                    choices=["cg_spmv", "dilu_fwd", "dilu_bwd"],
                    help="batch and piso modes: the kernel whose roofline is reported (the "
                         "channel reports all three, the dominant one first)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32-ir", "fp32-precon"],
                    help="pressure solve: FP64, FP32 inner Krylov in FP64 refinement, or "
                         "FP64 CG with the DILU sweeps in FP32")
    ap.add_argument("--inner-rtol", type=float, default=1e-6,
                    help="fp32-ir: relative tolerance of each FP32 inner solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick", action="store_true",
                    help="channel mode: skip the mixed-precision and batched summaries")
    ap.add_argument("--mode", default="channel",
                    choices=["channel", "batch", "serve", "stepbatch", "piso", "calib"],
                    help="channel: the 1M-cell FP64 channel; batch: batched cavities; "
                         "serve: Poisson-arrival serving of step cases; piso: the channel "
                         "as a transient PISO case")
    ap.add_argument("--requests", type=int, default=16, help="serve mode: number of requests")
    ap.add_argument("--serve-requests", type=int, default=8,
                    help="channel mode: requests of the serving summary (detail.serving)")
    ap.add_argument("--rate", type=float, default=2.0, help="serve mode: arrivals per second")
    ap.add_argument("--slots", type=int, default=4, help="serve mode: concurrent slots")
    ap.add_argument("--lanes", type=int, default=2,
                    help="serving: batches stepped concurrently (a case joins the lane "
                         "closest to it in size)")
    ap.add_argument("--hbm-gb", type=float, default=64.0, help="serve mode: HBM budget (GiB)")
    ap.add_argument("--max-iterations", type=int, default=1000,
                    help="serve mode: SIMPLE iterations per request at most")
    ap.add_argument("--precon", default="Jacobi", choices=["DILU", "Jacobi"],
                    help="serve mode: preconditioner of every solve (Jacobi, as the "
                         "reference's k-omega step example)")
    ap.add_argument("--no-alone", action="store_true",
                    help="serve mode: skip re-running each request alone")
    ap.add_argument("--laminar", action="store_true",
                    help="serve mode: laminar cases instead of k-omega RANS")
    ap.add_argument("--cases", type=int, default=256, help="batch mode: number of cases")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the path has no multi-GPU sharding); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    x = ge.load_package()
    if args.mode == "batch":
        run_batch(args, x)
        return
    if args.mode == "serve":
        run_serve(args, x)
        return
    if args.mode == "stepbatch":
        run_stepbatch(args, x)
        return
    if args.mode == "calib":
        run_calib(args)
        return
    if args.mode == "piso":
        run_piso(args, x)
        return
    api, H = x.api, x.hip
    N = args.cells
    t0 = time.perf_counter()
    mesh = x.mesh.box3d(N, N, N, jitter=0.1, seed_jitter=SEED_JITTER, seed_perm=SEED_PERM)
    t_mesh = time.perf_counter() - t0
    log(f"mesh built: {mesh.n_cells} cells in {t_mesh:.1f} s")
    model, config = channel(api, mesh, args.warmup + args.steps, args.precision, args.inner_rtol)
    setup = api.simple_setup(model, config)
    ctx = H.Context(0)
    t0 = time.perf_counter()
    dm = H.DeviceMesh(ctx, mesh)
    sim = H.Simple(dm, setup)
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    ctx.synchronize()
    t_upload = time.perf_counter() - t0
    log(f"uploaded in {t_upload:.1f} s")
    if args.warmup:
        sim.run(args.warmup)
        log(f"{args.warmup} warmup steps done")
    kernels = ("dilu_fwd", "dilu_bwd", "cg_spmv")
    for k in kernels:
        ctx.time_kernel(k, True)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = sim.run(args.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    timings = {k: ctx.kernel_time(k) for k in kernels}
    for k in kernels:
        ctx.time_kernel(k, False)
    if out["done"] != args.steps:
        raise RuntimeError(f"SIMPLE stopped after {out['done']} of {args.steps} steps")
    wall = t1 - t0
    log(f"{args.steps} timed steps in {wall:.2f} s")
    n = mesh.n_cells
    value = n * args.steps / wall
    stages = sim.stage_times()
    kit = out["krylov_iters"]
    vb = 8 if args.precision == "fp64" else 4

    def roof(k):
        launches, kms = timings[k]
        nbytes = kernel_bytes(k, n, dm.nnz, 8 if k == "cg_spmv" else vb,
                              8 if args.precision == "fp32-precon" else None)
        avg_s = kms / 1e3 / max(launches, 1)
        achieved = nbytes / avg_s / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": hbm_traffic(k, f"channel_{N}cubed_{args.precision}"),
                "kernel": k, "bytes_per_launch": nbytes, "timed_launches": launches,
                "avg_us": round(avg_s * 1e6, 2)}
    # the dominant kernel: most time per CG iteration (every timed kernel
    # launches once per iteration)
    dominant = max(kernels, key=lambda k: timings[k][1] / max(timings[k][0], 1))
    roofline = roof(dominant)
    roofline["others"] = [roof(k) for k in kernels if k != dominant]

    err = None
    if args.precision != "fp64":
        # per-field error of the mixed-precision run against the FP64 path
        # (which matches the CPU oracle to 1e-8 on the parity meshes)
        m64, c64 = channel(api, mesh, args.warmup + args.steps)
        sim64 = H.Simple(dm, api.simple_setup(m64, c64))
        sim64.set_U(m64.momentum.U)
        sim64.set_p(m64.momentum.p)
        sim64.run(args.warmup + args.steps)

        def rel(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        err = {f: rel(sim.get(f), sim64.get(f)) for f in ("U", "p", "mdotf")}

    mixed = batched = serving = None
    nnz, levels = dm.nnz, list(dm.levels)
    if args.precision == "fp64" and not args.quick:
        # the same channel with FP32 inner solves in FP64 refinement, and the
        # batched form: the other two parts of the headline metric
        def rel(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        def mixed_run(prec):
            mm, mc = channel(api, mesh, args.warmup + args.steps, prec, args.inner_rtol)
            simm = H.Simple(dm, api.simple_setup(mm, mc))
            simm.set_U(mm.momentum.U)
            simm.set_p(mm.momentum.p)
            if args.warmup:
                simm.run(args.warmup)
            ctx.synchronize()
            t0 = time.perf_counter()
            om = simm.run(args.steps)
            ctx.synchronize()
            wm = time.perf_counter() - t0
            return {"value": round(n * om["done"] / wm, 1), "unit": "cell-iters/s",
                    "ms_per_step": round(1e3 * wm / max(om["done"], 1), 3),
                    "p_iters_mean": float(np.mean(om["krylov_iters"][3])),
                    "error_vs_fp64": {f: rel(simm.get(f), sim.get(f))
                                      for f in ("U", "p", "mdotf")}}
        mixed = mixed_run("fp32-ir")
        mixed["p_solver"] = (f"FP32 CG+DILU (inner rtol {args.inner_rtol:g}) in FP64 "
                             f"refinement to 1e-10; p_iters_mean counts inner iterations")
        log(f"mixed precision (fp32-ir): {mixed['value']:.4g} cell-iters/s")
        pc = mixed_run("fp32-precon")
        pc["p_solver"] = "FP64 CG to 1e-10, DILU sweeps in FP32 (no restarts)"
        mixed["fp32_precon"] = pc
        log(f"mixed precision (fp32-precon): {pc['value']:.4g} cell-iters/s")
        batched = {}
        for prec in ("fp64", "fp32-ir"):
            bl = batch_line(args, x, prec, steps=10, warmup=2)
            batched[prec] = {"value": bl["value"], "unit": bl["unit"],
                             "case_iters_per_s": bl["detail"]["case_iters_per_s"],
                             "ms_per_step": bl["ms_per_step"], "cells": bl["config"]["cells"],
                             "roofline_frac": bl["roofline"]["frac"],
                             "graph": bl["detail"]["graph"], "tuned": bl["detail"]["tuned"]}
            log(f"batch {prec}: {bl['value']:.4g} cell-iters/s")
        batched["workload"] = bl["config"]["workload"].split(", SIMPLE")[0]
        # the server's lanes and submission path each take a stream; the
        # channel's context (one more) is released first, so that the
        # process holds no more streams than in --mode serve
        del sim, dm, ctx
        gc.collect()
        _, serving = serve_leg(args, x, args.serve_requests, args.rate, args.max_iterations,
                               alone=False)
        log(f"serving: {serving['converged']}/{args.serve_requests} converged, p50 "
            f"{serving['p50_time_to_converged_s']} s, p99 {serving['p99_time_to_converged_s']} s")

    cpu = c1 = None
    if not args.no_cpu and args.precision == "fp64":
        O = ge.load_oracle()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        c1 = config1(x, O, threads)
        log(f"config 1 (40x40 cavity): CPU {c1['cpu']['value']:.4g}, "
            f"GPU {c1['gpu']['value']:.4g} cell-iters/s")
        O.set_threads(threads)
        om = O.OracleMesh(mesh, x._abi)
        m2, c2 = channel(api, mesh, args.cpu_steps)
        s2 = api.simple_setup(m2, c2)
        t0 = time.perf_counter()
        ref = om.simple_run(s2, m2.momentum.U, m2.momentum.p, args.cpu_steps)
        tc = time.perf_counter() - t0
        cpu = {"value": round(n * ref["done"] / tc, 1), "unit": "cell-iters/s",
               "cores": threads, "kind": "port",
               "sample": f"first {ref['done']} SIMPLE iterations of the same 1M-cell FP64 "
                         f"channel, C++ oracle with OpenMP ({tc:.1f} s)"}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32-ir": "f64 (pressure Krylov f32 in f64 refinement)",
                  "fp32-precon": "f64 (pressure CG f64, DILU sweeps f32)"}[args.precision],
        "data": "synthetic (seeded jittered, renumbered hex box)",
        "config": {"workload": f"3D laminar channel SIMPLE {args.precision.upper()}, {N}x{N}x{N} "
                               f"jittered+renumbered hex box ({n} cells)",
                   "cells": n, "nnz": nnz, "dilu_levels": levels,
                   "p_solver": {"fp64": "CG+DILU rtol 1e-10",
                                "fp32-ir": f"FP32 CG+DILU (inner rtol {args.inner_rtol:g}) in "
                                           f"FP64 refinement to 1e-10",
                                "fp32-precon": "FP64 CG rtol 1e-10, DILU sweeps in FP32"}[
                                    args.precision],
                   "U_solver": "BiCGSTAB+DILU rtol 0.1",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {"mesh_build_s": round(t_mesh, 2), "upload_s": round(t_upload, 2),
                   "device_ms": {k: round(v, 2) for k, v in stages.items()},
                   "p_cg_iters_mean": float(np.mean(kit[3])),
                   "U_bicgstab_iters_mean": float(np.mean(kit[:3])),
                   "error_vs_fp64": err, "mixed_precision": mixed, "batched": batched,
                   "serving": serving, "config1": c1},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
