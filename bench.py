#!/usr/bin/env python
"""Benchmark of the fully-implicit oil-water TPFA Newton step on one MI355X.

Metric: Mcell-Newton-its/s = sum over Newton iterations of the cells in that
solve / wall seconds / 1e6, on the 60x220x85 field waterflood (1,122,000
cells, 10 report steps of 30 days).  A "step" is one report step of the
schedule: every Newton iteration in it (assembly, convergence check, CPR
setup, FGMRES, update) on inputs resident in HBM.  The mixed form (FP32 CPR
preconditioner inside FP64 FGMRES) is the headline value; the plain FP64
form runs the same schedule beside it and the maximum relative pressure
difference between the two is reported.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import os

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import argparse
import json
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# HBM bytes per k_bcsr_spmv<2,8> launch on the 60x220x85 Jacobian, from the
# FETCH_SIZE / WRITE_SIZE counter runs of `bench.py --traffic-probe 5`,
# calibrated against a 512 MiB copy (fetch factor 2.0, write factor 1.0):
# 404.3 MB read + 17.95 MB written.
MEASURED_SPMV_TRAFFIC = {(60, 220, 85): 422.2e6}
# k_line_gs_tw<2,float,8> per launch (one colour), mean over the ten launches
# of one preconditioner application, from the same counter runs over the
# first report step of the 60x220x85 run: 32.87 MB read + 8.92 MB written in
# each of the eight launches of the first four sweeps, 40.82 + 24.63 MB in
# each of the two of the last sweep.
MEASURED_LINE_GS_TRAFFIC = {(60, 220, 85): 46.5e6}


def spmv_bytes(nnzb: int, nc: int) -> float:
    """Algorithmic bytes of one 2x2-block BCSR SpMV launch: every block once
    (32 B of FP64 values + 4 B column index), the row pointers, and the cell
    parts of x and y once each (16 B per cell each)."""
    return nnzb * 36.0 + 4.0 * (nc + 1) + 16.0 * nc * 2


# CPR settings of the oil-water runs (field case and ensemble, both forms of
# each comparison): five line Gauss-Seidel sweeps instead of the default four
# (the black-oil case runs best at four); with undamped Jacobi smoothing the
# field case takes 1,480 FGMRES iterations instead of 1,796 at the earlier
# 4 sweeps / omega 0.67, with the same Newton iterations
OW_CPR = {"local_sweeps": 5, "jacobi_omega": 1.0}
# the black-oil case (200k cells, 20-cell lines): three sweeps, both forms
# (3,895 -> 4,065 FGMRES iterations but 2-4% faster than four; same Newton)
BO_CPR = {"local_sweeps": 3}


def run_form(case, precision, steps, warmup, rtol, local_stage=None, line_segments=False):
    """Time `steps` report steps of the schedule: passes over it from the
    initial state, the last one cut short, so any step count is exact."""
    import torch
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import build_model
    cpr = dict(OW_CPR)
    if local_stage is not None:
        cpr["local_stage"] = local_stage
    if line_segments:
        cpr["line_segments"] = True
    # no per-solve log lines inside the timed region (the observer reports each report step)
    solver = CPRSolver(rtol=rtol, precision=precision, verbose=0, **cpr)
    t_setup = time.perf_counter()
    model, s0 = build_model(case, solver)
    print(f"[bench] FP{precision}: model set up in {time.perf_counter() - t_setup:.1f} s", file=sys.stderr,
          flush=True)
    dts = case.dts

    def progress(event, payload):
        if event in ("step_accepted", "step_rejected", "timestep_cut"):
            print(f"[bench] FP{precision} {event} {payload}", file=sys.stderr, flush=True)

    if warmup > 0:
        simulate_schedule_ad(s0, model, dts[:min(warmup, dts.size)], keep_device=True, observer=progress)
    model.reset()
    solver.history.clear()
    model.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    states, its, done = [], [], 0
    while done < steps:
        n = min(steps - done, dts.size)
        if done:
            model.reset()          # a new pass over the schedule starts from the initial state
        st, it = simulate_schedule_ad(s0, model, dts[:n], keep_device=True, observer=progress)
        states += st
        its.append(it)
        done += n
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    its = np.concatenate(its)
    prof = model.profile_read()
    model.profile(False)
    p_hist = [s.p.cpu().numpy() for s in states]
    return dict(model=model, its=its, wall=wall, p_hist=p_hist, prof=prof,
                lin_its=sum(h[0] for h in solver.history), n_solves=len(solver.history))


def incomp_leg(reps: int = 5):
    """The incompTPFA pressure solve (32x32x8, lognormal permeability,
    gravity, rate injector and bhp producer columns): device assembly, AMG-
    FGMRES solve and reconstruction timed over `reps` solves, against the
    oracle's assembly with SciPy's SuperLU solve (the reference's
    spsolve, incomp_tpfa.py:172-184), single thread, on the same inputs."""
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle.grid as og
    from oracle.incomp import incomp_tpfa as o_incomp
    from pymrst_amd import incomp as pi
    from pymrst_amd.workloads import incomp_case
    c = incomp_case()
    g, rock, wells, fluid, gv = c["grid"], c["rock"], c["wells"], c["fluid"], c["gravity"]
    solver = pi.IncompTPFA(g)
    T = solver.compute_trans(rock)
    st0 = pi.init_res_sol(g, 0.0)
    solver.solve(st0, T, fluid, wells=wells, gravity=gv)          # warm-up: allocations, AMG hierarchy
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = solver.solve(st0, T, fluid, wells=wells, gravity=gv)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    go = og.cart_grid(c["dims"], c["phys"])
    ro = og.make_rock(go, c["perm"], c["poro"])
    wo = None
    for cells, kind, val in c["well_specs"]:
        wo = og.add_well(wo, go, ro, cells, kind, val, compi=[1.0], gravity=gv)
    to = og.compute_trans(go, ro)
    mob = np.full((go.num_cells, 1), 1.0 / c["mu"])
    times = []
    for _ in range(3):
        t1 = time.perf_counter()
        ref = o_incomp(go, to, mob, np.array([c["rho"]]), wells=wo, gravity=gv)
        times.append(time.perf_counter() - t1)
    cpu_s = float(np.median(times))
    err = float(np.abs(st.pressure - ref.pressure).max() / np.abs(ref.pressure).max())
    nc = g.num_cells
    return {"workload": f"incompTPFA {c['dims'][0]}x{c['dims'][1]}x{c['dims'][2]} ({nc} cells), lognormal k, "
                        "gravity, rate injector + bhp producer columns",
            "gpu_ms_per_solve": round(gpu_s * 1e3, 3), "fgmres_its": int(solver.last_info.iterations),
            "max_rel_pressure_diff_vs_oracle": err,
            "cpu_baseline": {"value": round(cpu_s * 1e3, 3), "unit": "ms per solve", "cores": 1, "kind": "port",
                             "sample": "the oracle's incompTPFA (assembly, SciPy SuperLU spsolve, flux "
                                       "reconstruction) on the same inputs, median of 3, 1 thread"}}


def cpu_baseline(case_sample, budget_s=25.0):
    """The oracle (NumPy/SciPy restatement of the reference's Newton + CPR,
    single thread) on a bounded sample: Newton iterations of the first report
    step of a smaller grid of the same workload family."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.cases import model_from_case
    from oracle.cpr import CPRSolverCPU
    c = case_sample
    m, s0 = model_from_case(c, CPRSolverCPU())
    g = m.g
    st = s0.copy()
    cdp = m.compute_cdp_per_well(s0)
    dt = float(c.dts[0])
    its = 0
    t0 = time.perf_counter()
    while its < m.max_iterations and time.perf_counter() - t0 < budget_s:
        r, J, eqs = m.residual_and_jacobian(s0, st, dt, cdp)
        if m.converged(eqs, dt, st)[0]:
            break
        m.apply_update(st, m.solve_increment(r, J, st))
        its += 1
    wall = time.perf_counter() - t0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return dict(value=g.num_cells * its / wall / 1e6, unit="Mcell-Newton-its/s", cores=1, kind="port",
                sample=(f"oracle Newton+CPR(spilu pressure, block-Jacobi), first {its} Newton iterations of "
                        f"report step 1 on {c.dims[0]}x{c.dims[1]}x{c.dims[2]} ({g.num_cells} cells) of the "
                        f"same workload; {wall:.1f} s, 1 thread of {cores} available"))


def bo_spmv_bytes(nnzb: int, nc: int) -> float:
    """Algorithmic bytes of one 3x3-block BCSR SpMV launch: every block once
    (72 B of FP64 values + 4 B column index), the row pointers, and the cell
    parts of x and y once each (24 B per cell each)."""
    return nnzb * 76.0 + 4.0 * (nc + 1) + 24.0 * nc * 2


def bo_cpu_baseline(dims=(30, 30, 10), budget_s=12.0):
    """The oracle black-oil Newton (forward-mode AD assembly, direct solve,
    update and flash; single thread) on the first report step of a smaller
    grid of the same workload family."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import scipy.sparse.linalg as spla
    from oracle.cases import bo_model_from_case
    from pymrst_amd.grid import DAY
    from pymrst_amd.workloads import blackoil_field
    c = blackoil_field(dims=dims)
    m, s0 = bo_model_from_case(c)
    dt = float(c.dts[0]) / 4.0
    st, its = s0.copy(), 0
    t0 = time.perf_counter()
    while its < m.max_iterations and time.perf_counter() - t0 < budget_s:
        r, J, eqs = m.residual_and_jacobian(s0, st, dt)
        if m.converged(eqs, dt, st)[0]:
            break
        st = m.update_state(st, spla.spsolve(J.tocsc(), -r))
        its += 1
    wall = time.perf_counter() - t0
    nc = m.nc
    return dict(value=nc * its / wall / 1e6, unit="Mcell-Newton-its/s", cores=1, kind="port",
                sample=(f"oracle black-oil Newton (AD assembly, scipy direct solve, flash), first {its} Newton "
                        f"iterations of a {dt / DAY:g}-day step on {dims[0]}x{dims[1]}x{dims[2]} ({nc} cells) of "
                        f"the same workload; {wall:.1f} s, 1 thread"))


def blackoil_leg(steps: int = 20, cpu: bool = True):
    """The three-phase live-oil case (100x100x20, gas cap injector, four
    producers, `steps` steps of 15 days) with the FP64 and the FP32
    preconditioner: whole-schedule Mcell-Newton-its/s (cuts included),
    Newton and GMRES counts, the mixed form's pressure error against FP64,
    and the 3x3 BCSR SpMV's achieved bandwidth."""
    import torch
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import blackoil_field, build_bo_model
    case = blackoil_field(steps=steps)
    nc = case.num_cells
    out, p_last = {}, {}
    for prec in (64, 32):
        solver = CPRSolver(rtol=1e-8, precision=prec, **BO_CPR)
        model, s0 = build_bo_model(case, solver)
        cuts = [0]

        def progress(event, payload):
            if event == "timestep_cut":
                cuts[0] += 1
            if event in ("step_accepted", "timestep_cut"):
                print(f"[bench] black-oil FP{prec} {event} {payload}", file=sys.stderr, flush=True)

        model.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        states, its = simulate_schedule_ad(s0, model, case.dts, keep_device=True, observer=progress)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = model.profile_read()
        model.profile(False)
        newton = int(model.stats["newton"])
        nbytes = bo_spmv_bytes(model.nnzb, nc)
        avg_s = prof["spmv_ms"] / max(prof["spmv_count"], 1) / 1e3
        h = states[-1].to_host()
        p_last[prec] = h.pressure
        out[f"fp{prec}"] = dict(value=round(nc * newton / wall / 1e6, 3), wall_s=round(wall, 3),
                                newton_its=newton, newton_its_per_step=its.tolist(), cuts=cuts[0],
                                gmres_its=int(sum(x[0] for x in solver.history)),
                                spmv={"achieved_GBs": round(nbytes / avg_s / 1e9, 1) if prof["spmv_count"] else None,
                                      "avg_launch_us": round(avg_s * 1e6, 2), "launches": prof["spmv_count"],
                                      "bytes_per_launch": nbytes},
                                saturated_cells_end=int((h.status == 3).sum()),
                                launch_shapes=model.launch_shapes())
        print(f"[bench] black-oil FP{prec}: {newton} Newton its, {wall:.2f} s", file=sys.stderr, flush=True)
        del model
    out["workload"] = (f"black-oil {case.dims[0]}x{case.dims[1]}x{case.dims[2]} ({nc} cells), gas-cap injector "
                       f"1e6 Sm3/d, 4 producers at 150 bar, {steps} x 15 d, CPR-FGMRES rtol 1e-8")
    out["max_rel_pressure_diff_mixed_vs_fp64"] = float(np.max(np.abs(p_last[32] - p_last[64]) /
                                                              np.abs(p_last[64])))
    out["speedup_mixed_vs_fp64"] = round(out["fp32"]["value"] / out["fp64"]["value"], 3)
    out["cpr"] = dict(BO_CPR)
    if cpu:
        out["cpu_baseline"] = bo_cpu_baseline()
    return out


def adjoint_leg(fd_cells: int = 0, seed: int = 5):
    """The log-permeability adjoint on the 64x64x16 oil-water case (20 steps
    of 15 days): one forward run keeping checkpoints, one backward sweep,
    optionally central finite differences (h = 1e-6) on `fd_cells` seeded
    random cells."""
    import torch
    from pymrst_amd.workloads import adjoint_problem
    prob, logk = adjoint_problem()
    prob.forward(logk)          # warm-up: library, allocations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run = prob.forward(logk)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g, gu = prob.gradient(run)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    J = prob.value(run)
    out = {"workload": "oil-water 64x64x16 (65,536 cells), 20 x 15 d, d J / d ln k for every cell",
           "objective": J, "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
           "gradients_per_s": round(1.0 / (t2 - t0), 3), "newton_its": int(sum(run.newton)),
           "adjoint_fgmres_its": int(sum(prob.adjoint_its))}
    if fd_cells > 0:
        rng = np.random.default_rng(seed)
        cells = rng.choice(logk.size, size=fd_cells, replace=False)
        gh = g.cpu().numpy()
        fd = np.zeros(fd_cells)
        h = 1e-6
        for i, c in enumerate(cells):
            lp, lm = logk.copy(), logk.copy()
            lp[c] += h
            lm[c] -= h
            fd[i] = (prob.value(prob.forward(lp)) - prob.value(prob.forward(lm))) / (2 * h)
            print(f"[bench] adjoint FD cell {int(c)}: adjoint {gh[c]:.9e} FD {fd[i]:.9e}", file=sys.stderr,
                  flush=True)
        adj = gh[cells]
        out["fd_check"] = {"cells": fd_cells, "h": h,
                           "max_abs_err_over_max_fd": float(np.abs(adj - fd).max() / np.abs(fd).max()),
                           "max_rel_err_per_cell": float((np.abs(adj - fd) / np.abs(fd)).max()),
                           "median_rel_err_per_cell": float(np.median(np.abs(adj - fd) / np.abs(fd)))}
    return out


def ensemble_leg(members: int = 256, loop_sample: int = 16, warmup: int = 1):
    """The uneven 256-member ensemble (5 x 30 d): the batched form (one
    union system, FP32 CPR inside FP64 FGMRES) against the plain FP64 form
    looped member by member over the first `loop_sample` members (a time
    box; the per-member rate is what the metric compares)."""
    import torch
    from pymrst_amd.ensemble import build_ensemble, simulate_ensemble
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import build_model, ensemble_cases
    cases = ensemble_cases(members)
    dts = cases[0].dts
    # a Newton path that diverges for one member makes its linear solve
    # fail; 20 restart cycles let that fail fast into a time-step cut
    solver = CPRSolver(precision=32, verbose=0, maxiter=20, **OW_CPR)
    t0 = time.perf_counter()
    model, s0, ranges = build_ensemble(cases, solver)
    model.max_iterations = 40
    build_s = time.perf_counter() - t0
    print(f"[bench] ensemble: union of {members} members, {model.nc} cells, {model.nw} wells, built in "
          f"{build_s:.1f} s", file=sys.stderr, flush=True)

    def progress(event, payload):
        if event in ("step_accepted", "step_rejected"):
            print(f"[bench] ensemble {event} {payload}", file=sys.stderr, flush=True)

    # per-member time stepping: every member cuts and regrows on its own
    if warmup:
        simulate_ensemble(model, s0, dts[:warmup], keep_device=True, observer=progress)
    model.reset()
    solver.history.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    states, its, member_its = simulate_ensemble(model, s0, dts, keep_device=True, observer=progress)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ncells = model.nc
    member_cells = np.array([b - a for a, b in ranges])
    per_member = member_its.sum(axis=0)
    useful = int((member_cells * per_member).sum())
    batched = {"value": round(useful / wall / 1e6, 3), "wall_s": round(wall, 3),
               "newton_its_union": its.tolist(),
               "newton_its_per_member_mean": round(float(per_member.mean()), 2),
               "gmres_its": int(sum(h[0] for h in solver.history)), "union_build_s": round(build_s, 1)}
    model_its_sample = per_member[:loop_sample].copy()
    del model, states
    torch.cuda.empty_cache()
    tot_cells_its, tot_wall, ncl, loop_its = 0, 0.0, 0, []
    for k, c in enumerate(cases[:loop_sample]):
        print(f"[bench] ensemble: looped member {k} ({c.num_cells} cells)", file=sys.stderr, flush=True)
        sv = CPRSolver(precision=64, maxiter=20, **dict(OW_CPR, autotune=False))
        m, st0 = build_model(c, sv)
        m.max_iterations = 40
        if warmup:
            simulate_schedule_ad(st0, m, c.dts[:warmup], keep_device=True)
        m.reset()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, its_m = simulate_schedule_ad(st0, m, c.dts, keep_device=True)
        torch.cuda.synchronize()
        tot_wall += time.perf_counter() - t1
        tot_cells_its += c.num_cells * int(its_m.sum())
        ncl += c.num_cells
        loop_its.append(int(its_m.sum()))
        del m
    looped = {"members": loop_sample, "cells": ncl, "value": round(tot_cells_its / tot_wall / 1e6, 3),
              "wall_s": round(tot_wall, 3), "newton_its_per_member_mean": round(float(np.mean(loop_its)), 2),
              "us_per_member_cell": round(tot_wall / ncl * 1e6, 3)}
    batched["us_per_member_cell"] = round(wall / ncells * 1e6, 3)
    batched["newton_its_same_members_mean"] = round(float(model_its_sample.mean()), 2)
    print(f"[bench] ensemble: batched {batched['value']} vs looped FP64 {looped['value']} Mcell-Newton-its/s",
          file=sys.stderr, flush=True)
    return {"workload": f"{members} oil-water realisations, nx,ny~U{{20..64}}, nz~U{{5..16}}, "
                        f"sigma(ln k)~U(0.5,2), 5 x 30 d",
            "metric_note": "cell-Newton-iterations of each member (each with its own time steps; members "
                           "that have converged or reached the report time are frozen) per wall second",
            "members": members, "cells": ncells, "cpr": dict(OW_CPR),
            "batched_fp32_precond": batched, "looped_plain_fp64_sample": looped,
            "speedup_batched_vs_looped_fp64": round(batched["value"] / looped["value"], 2),
            "time_to_solution_ratio_looped_vs_batched": round(looped["us_per_member_cell"] /
                                                              batched["us_per_member_cell"], 2)}


def stream_copy(reps: int = 20, mib: int = 1024):
    """STREAM-style copy bandwidth of this GPU: `reps` device-to-device
    copies of a `mib` MiB FP64 buffer (four times the 256 MB Infinity
    Cache), timed with events on torch's stream; read and write bytes both
    counted."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * n * 8
    del a, b
    return {"GBs": round(moved / ms / 1e6, 1), "bytes_per_copy": moved, "avg_ms": round(ms, 4),
            "buffer_MiB": mib, "reps": reps}


def traffic_probe(case, reps):
    """A short program for hardware-counter runs: the FP64 BCSR SpMV of the
    assembled Jacobian `reps` times, then a plain 512 MiB device copy `reps`
    times (the copy's known bytes calibrate the FETCH_SIZE / WRITE_SIZE
    counters)."""
    import ctypes
    import torch
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    from pymrst_amd.workloads import build_model
    model, s0 = build_model(case, CPRSolver(precision=32))
    d0 = DeviceOWState.from_host(s0, model.device)
    cdp = model.compute_cdp_per_well(d0)
    model.assemble_device(d0, d0, float(case.dts[0]), cdp)
    n = model._dx.numel()
    x = torch.randn(n, dtype=torch.float64, device=model.device)
    y = torch.empty_like(x)
    for _ in range(reps):
        _lib.check(model.lib.pmx_ow_matvec(model.h, _lib.ptr(x), _lib.ptr(y)), "pmx_ow_matvec")
    a = torch.ones(1 << 26, dtype=torch.float64, device=model.device)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "spmv+copy", "reps": reps, "spmv_bytes": spmv_bytes(model.nnzb, case.num_cells),
                      "copy_read_bytes": a.numel() * 8, "copy_write_bytes": b.numel() * 8}))
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--form", choices=("both", "mixed", "fp64"), default="both")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--dims", default="60,220,85")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-fd", type=int, default=0, metavar="CELLS",
                    help="check the adjoint gradient by central differences on CELLS random cells")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-blackoil", action="store_true")
    ap.add_argument("--blackoil-only", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-loop", type=int, default=16, help="members timed in the looped plain form")
    ap.add_argument("--local-stage", choices=("line", "ilu", "block_jacobi"), default=None,
                    help="CPR local stage of the oil-water legs (default: line Gauss-Seidel)")
    ap.add_argument("--line-segments", action="store_true",
                    help="solve lines of 32+ cells by segments with spikes (partitioned) instead of one walk")
    ap.add_argument("--autotune", type=int, choices=(0, 1), default=1,
                    help="1 (default): launch shapes chosen by timing at the first setup of each size class "
                         "(AMG level layouts, folded launches, SpMV lanes); 0: fixed shapes")
    ap.add_argument("--tune-file", default=None, metavar="PATH",
                    help="launch-shape table: read at the start when the file exists (those shapes are used "
                         "without timing them), written at the end")
    ap.add_argument("--no-incomp", action="store_true")
    ap.add_argument("--incomp-only", action="store_true")
    ap.add_argument("--traffic-probe", type=int, default=0, metavar="REPS",
                    help="run only the SpMV and a calibration copy REPS times (for counter runs)")
    args = ap.parse_args()
    OW_CPR["autotune"] = BO_CPR["autotune"] = bool(args.autotune)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "realisation or by y-slabs, which this build does not run across GPUs", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if args.tune_file:
        import atexit
        from pymrst_amd import _lib
        if os.path.exists(args.tune_file):
            n = _lib.autotune_load(args.tune_file)
            print(f"[bench] launch shapes: {n} entries read from {args.tune_file}", file=sys.stderr, flush=True)
        atexit.register(_lib.autotune_save, args.tune_file)
    from pymrst_amd.workloads import field_waterflood
    dims = tuple(int(v) for v in args.dims.split(","))
    case = field_waterflood(dims=dims)
    if args.traffic_probe > 0:
        return traffic_probe(case, args.traffic_probe)
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(args.ensemble_members, args.ensemble_loop)}))
        return 0
    if args.blackoil_only:
        print(json.dumps({"blackoil": blackoil_leg(cpu=not args.no_cpu_baseline)}))
        return 0
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(args.adjoint_fd)}))
        return 0
    if args.incomp_only:
        print(json.dumps({"incomp": incomp_leg()}))
        return 0
    steps = max(1, args.steps)
    nc = case.num_cells
    forms = {"both": (64, 32), "mixed": (32,), "fp64": (64,)}[args.form]
    res = {}
    for prec in forms:
        res[prec] = run_form(case, prec, steps, args.warmup, args.rtol, args.local_stage, args.line_segments)
        r = res[prec]
        print(f"[bench] FP{prec} preconditioner: {int(r['its'].sum())} Newton its, {r['lin_its']} GMRES its, "
              f"{r['wall']:.3f} s", file=sys.stderr, flush=True)
    head = res[32] if 32 in res else res[64]
    newton = int(head["its"].sum())
    value = nc * newton / head["wall"] / 1e6
    extra = {}
    if 32 in res and 64 in res:
        p64 = np.vstack(res[64]["p_hist"])
        p32 = np.vstack(res[32]["p_hist"])
        extra["max_rel_pressure_diff_mixed_vs_fp64"] = float(np.max(np.abs(p32 - p64) / np.abs(p64)))
        extra["fp64_value"] = nc * int(res[64]["its"].sum()) / res[64]["wall"] / 1e6
        extra["newton_its_fp64"] = res[64]["its"].tolist()
        extra["speedup_mixed_vs_fp64"] = value / extra["fp64_value"]
    extra["newton_its"] = head["its"].tolist()
    extra["launch_shapes"] = head["model"].launch_shapes()
    extra["gmres_its"] = head["lin_its"]
    prof = head["prof"]
    m = head["model"]
    nbytes = spmv_bytes(m.nnzb, nc)
    roof = None
    if prof["spmv_count"] > 0:
        avg_s = prof["spmv_ms"] / prof["spmv_count"] / 1e3
        achieved = nbytes / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": MEASURED_SPMV_TRAFFIC.get(dims),
                "kernel": "k_bcsr_spmv<2,8> (FP64 2x2 BCSR SpMV of the Jacobian)",
                "bytes_per_launch": nbytes, "avg_launch_us": round(avg_s * 1e6, 2),
                "launches": prof["spmv_count"]}
    # the kernel with the most GPU time is the line solve of the CPR local
    # stage; it runs inside the replayed preconditioner graph, so it is timed
    # launch by launch in eager applications on the last Newton system
    roof_spmv, roof = roof, None
    if head["model"].linear_solver.local_stage == "line" and nc > 100000:
        lg = head["model"].time_line_solve(20)
        if lg["launches"] > 0:
            avg_s = lg["ms"] / lg["launches"] / 1e3
            per_launch = lg["bytes"] / lg["launches"]
            achieved = per_launch / avg_s / 1e9
            roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": MEASURED_LINE_GS_TRAFFIC.get(dims),
                    "kernel": ("k_line_gs_seg<2,float,8,4> (partitioned" if args.line_segments else
                               "k_line_gs_tw<2,float,8> (twisted") +
                              " block-tridiagonal line solves of one colour, the CPR local stage; the kernel with "
                              "the most GPU time)",
                    "bytes_per_launch": round(per_launch), "avg_launch_us": round(avg_s * 1e6, 2),
                    "launches": lg["launches"]}
    if roof is None:
        roof, roof_spmv = roof_spmv, None
    cpu = None
    if not args.no_cpu_baseline:
        sample = field_waterflood(dims=(dims[0], max(2, dims[1] // 10), dims[2]))
        cpu = cpu_baseline(sample)
    line = {"metric": "Mcell-Newton-its/s", "value": round(value, 3), "unit": "Mcell-Newton-its/s",
            "n_gpus": 1, "steps": steps, "steps_requested": args.steps, "warmup": args.warmup,
            "ms_per_step": round(head["wall"] / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64 (FP32 preconditioner)" if 32 in res else "f64",
            "data": "synthetic",
            "config": {"workload": f"oil-water field waterflood {dims[0]}x{dims[1]}x{dims[2]} "
                                   f"({nc} cells), five-spot, gravity, {steps} report steps of 30 d (passes over "
                                   f"the {case.dts.size}-step schedule from the initial state), CPR-FGMRES rtol "
                                   f"{args.rtol}",
                       "cells": nc, "report_steps": steps,
                       "cpr": dict(OW_CPR, **({"local_stage": args.local_stage} if args.local_stage else {}),
                                   **({"line_segments": True} if args.line_segments else {}))},
            "roofline": roof, "roofline_spmv": roof_spmv, "cpu_baseline": cpu}
    line.update(extra)
    line["stream_copy"] = stream_copy()
    if not args.no_incomp:
        line["incomp"] = incomp_leg()
    if not args.no_blackoil:
        line["blackoil"] = blackoil_leg(cpu=not args.no_cpu_baseline)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(args.adjoint_fd)
    if not args.no_ensemble:
        line["ensemble"] = ensemble_leg(args.ensemble_members, args.ensemble_loop)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
