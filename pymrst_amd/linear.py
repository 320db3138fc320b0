"""The reference's linear-solver seam on the device: CPRSolver.solve(jac, rhs,
model, state, dt) for a Jacobian assembled by the caller.

Reference: pymrst/modules/ad_core/solvers/linear_solvers.py (CPRSolver.solve
:426-545, _weighting :155-190, _unmet_requirements :192-224); callers
NonLinearSolver._solve_linear (simulators/nonlinear_solver.py:383-423) and
TwoPhaseOilWaterModelAD._solve_increment (two_phase_oil_water_model.py
:597-603).

The reference orders unknowns variable-blocked: [v_0 (nc), .., v_{B-1} (nc),
q_0 (nw), .., q_{B-1} (nw), bhp (nw)] with B = model.n_phases, and the
equations the same way (phase equations per cell, then per-phase well
equations and the control equation per well).  The device system is
cell-major: B x B blocks per cell pair (any pattern: the wellbore mixture's
couplings between perforated cells become ordinary blocks), then B + 1
unknowns per well, coupled to the cells through per-perforation blocks.
A model that lacks what the pressure stage needs gets the reference's loud
fallback to a direct solve (linear_solvers.py:426-461), run on the device
(dense LU, up to DIRECT_MAX unknowns) instead of on the host.  The AD-OO
seam solve_linear_problem(problem, model) -> (dx, result, report)
(linear_solver_ad.py:832-893) and the elliptic-solver injection
(linear_solvers.py:124-131, 253-256) sit on the same device solve.
"""
from __future__ import annotations

import ctypes
import numpy as np
import scipy.sparse as sp

from . import _lib


def reference_to_device(nc: int, B: int, nw: int) -> np.ndarray:
    """dev[i]: the device index of the reference's unknown (or equation) i."""
    M = B + 1
    dev = np.empty(B * nc + M * nw, dtype=np.int64)
    cells = np.arange(nc, dtype=np.int64)
    for a in range(B):
        dev[a * nc + cells] = cells * B + a
    wells = np.arange(nw, dtype=np.int64)
    for m in range(M):
        dev[B * nc + m * nw + wells] = nc * B + wells * M + m
    return dev


def unmet_requirements(model) -> list:
    """What the pressure stage needs from the model (the reference's list)."""
    if model is None:
        return ["no model was supplied"]
    missing = []
    if not hasattr(model, "G"):
        missing.append("model has no .G")
    if not hasattr(model, "n_phases"):
        missing.append("model has no .n_phases")
    fluid = getattr(model, "fluid", None)
    if fluid is None:
        missing.append("model has no .fluid")
    else:
        for attribute in ("disgas", "vapoil", "b"):
            if not hasattr(fluid, attribute):
                missing.append(f"model.fluid has no .{attribute}")
    return missing


def pressure_weights(model, state, nph: int) -> np.ndarray:
    """The quasi-IMPES weights 1/b_i at the mean pressure (and mean rs / rv
    where the fluid dissolves gas / vaporises oil), as _weighting."""
    def mean(name):
        v = state[name] if isinstance(state, dict) else getattr(state, name)
        return float(np.mean(np.asarray(getattr(v, "val", v), dtype=float)))
    fluid = model.fluid
    p_mean = mean("pressure")
    rs = mean("rs") if fluid.disgas else None
    sat = np.ones(1, dtype=bool) if fluid.disgas else None
    rv = mean("rv") if fluid.vapoil else None
    oil_present = np.ones(1, dtype=bool) if fluid.vapoil else None
    b = fluid.b(np.array([p_mean]), rs, sat, rv, oil_present)
    return np.array([1.0 / float(np.asarray(getattr(bi, "val", bi)).ravel()[0]) for bi in b[:nph]])


class DeviceSystem:
    """A pmx_sys handle: the pattern (cells and perforations) fixed at
    construction, values and solves per Newton iteration; `keys` / `pkeys`
    hold the pattern as SplitSystem does."""

    def __init__(self, B: int, nc: int, rowptr, col, perf_ptr, perf_cells, device=None):
        import torch
        self.lib = _lib.load()
        self.device = torch.device(device or "cuda")
        self.B, self.M, self.nc = B, B + 1, nc
        self.nw = len(perf_ptr) - 1
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.perf_ptr = np.ascontiguousarray(perf_ptr, dtype=np.int32)
        self.perf_cells = np.ascontiguousarray(perf_cells, dtype=np.int32)
        h = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        P = _lib.ptr
        _lib.check(self.lib.pmx_sys_create(self.device.index or 0, ctypes.c_void_p(stream), B, nc, P(self.rowptr),
                                           P(self.col), self.nw, P(self.perf_ptr), P(self.perf_cells),
                                           ctypes.byref(h)), "pmx_sys_create")
        self.h = h

    def set_values(self, val, Bp, Cp, Dp):
        import torch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1), device=self.device)
        self._vals = [t(val), t(Bp) if self.nw else None, t(Cp) if self.nw else None, t(Dp) if self.nw else None]
        P = lambda x: _lib.ptr(x) if x is not None else None
        _lib.check(self.lib.pmx_sys_set_values(self.h, *(P(x) for x in self._vals)), "pmx_sys_set_values")

    def matvec(self, x):
        import torch
        y = torch.empty_like(x)
        _lib.check(self.lib.pmx_sys_matvec(self.h, _lib.ptr(x), _lib.ptr(y)), "pmx_sys_matvec")
        return y

    def set_pressure_solver(self, fn):
        """fn(b: ndarray) -> x: the pressure stage of every preconditioner
        application (None: back to the device AMG)."""
        if fn is None:
            self._pcb = None
            _lib.check(self.lib.pmx_sys_set_pressure_solver(self.h, None, None), "pmx_sys_set_pressure_solver")
            return
        errors = []

        def call(_user, n, b, x):
            try:
                bv = np.ctypeslib.as_array(b, shape=(n,)).copy()
                np.ctypeslib.as_array(x, shape=(n,))[:] = np.asarray(fn(bv), dtype=float).ravel()
                return 0
            except Exception as e:          # reported as a failed solve
                errors.append(e)
                return 1

        proto = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double))
        self._pcb = (proto(call), errors)       # kept alive while the library holds it
        _lib.check(self.lib.pmx_sys_set_pressure_solver(self.h, ctypes.cast(self._pcb[0], ctypes.c_void_p), None),
                   "pmx_sys_set_pressure_solver")

    def solve(self, opts, w, rhs):
        import torch
        x = torch.zeros_like(rhs)
        info = _lib.SolveInfo()
        wv = np.ascontiguousarray(w, dtype=np.float64)
        rc = self.lib.pmx_sys_solve(self.h, ctypes.byref(opts), _lib.ptr(wv), _lib.ptr(rhs), _lib.ptr(x),
                                    ctypes.byref(info))
        pcb = getattr(self, "_pcb", None)
        if rc < 0 and pcb is not None and pcb[1]:
            raise pcb[1][-1]
        if rc < 0:
            _lib.check(rc, "pmx_sys_solve")
        return x, rc, info

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and getattr(self, "lib", None) is not None:
            self.lib.pmx_sys_destroy(h)
            self.h = None


class SplitSystem(tuple):
    """(rowptr, col, val, perf_ptr, perf_cells, Bp, Cp, Dp, dev) with the
    joined block and perforation patterns as .keys and .pkeys."""

    def __new__(cls, *items):
        t = super().__new__(cls, items[:9])
        t.keys, t.pkeys = items[9], items[10]
        return t


def split_reference_jacobian(jac, nc: int, B: int, keys_hint=None, pkeys_hint=None):
    """The reference's variable-blocked Jacobian as device pieces: (rowptr,
    col, val[nnzb, B, B], perf_ptr, perf_cells, Bp, Cp, Dp, dev).  The block
    pattern is the matrix's own (every diagonal block added) joined with
    `keys_hint` (row * nc + col of blocks to keep), the perforations with
    `pkeys_hint` (well * nc + cell): the Jacobian's pattern follows the
    state (a mixture coupling exists only while a connection injects), and
    a pattern that only grows lets one device system serve the Newton loop.
    The joined patterns are returned as `.keys` / `.pkeys` of the result."""
    n = jac.shape[0]
    M = B + 1
    nw, rem = divmod(n - B * nc, M)
    if jac.shape != (n, n) or rem != 0 or nw < 0:
        raise ValueError(f"a {jac.shape} Jacobian is not {B} phases on {nc} cells plus {M} unknowns per well")
    dev = reference_to_device(nc, B, nw)
    a = sp.coo_matrix(jac)
    r, c, v = dev[a.row], dev[a.col], np.asarray(a.data, dtype=float)
    ncB = nc * B
    # cell blocks (every diagonal block present)
    m = (r < ncB) & (c < ncB)
    br, bc = r[m] // B, c[m] // B
    key = br * nc + bc
    parts = [key, np.arange(nc, dtype=np.int64) * (nc + 1)]
    if keys_hint is not None:
        parts.append(np.asarray(keys_hint, dtype=np.int64))
    keys = np.unique(np.concatenate(parts))
    rowb = keys // nc
    rowptr = np.searchsorted(rowb, np.arange(nc + 1)).astype(np.int32)
    col = (keys % nc).astype(np.int32)
    val = np.zeros((keys.size, B, B))
    np.add.at(val, (np.searchsorted(keys, key), r[m] % B, c[m] % B), v[m])
    # perforations: cells coupled to a well in either direction
    mcw = (r < ncB) & (c >= ncB)
    mwc = (r >= ncB) & (c < ncB)
    mww = (r >= ncB) & (c >= ncB)
    w_cw, c_cw = (c[mcw] - ncB) // M, r[mcw] // B
    w_wc, c_wc = (r[mwc] - ncB) // M, c[mwc] // B
    pparts = [w_cw * nc + c_cw, w_wc * nc + c_wc]
    if pkeys_hint is not None:
        pparts.append(np.asarray(pkeys_hint, dtype=np.int64))
    pk = np.unique(np.concatenate(pparts))
    pw = pk // nc
    perf_ptr = np.searchsorted(pw, np.arange(nw + 1)).astype(np.int32)
    perf_cells = (pk % nc).astype(np.int32)
    if nw and np.any(np.diff(perf_ptr) == 0):
        raise ValueError("a well without any coupling to a cell")
    nperf = pk.size
    Bp = np.zeros((max(nperf, 1), B, M))
    Cp = np.zeros((max(nperf, 1), M, B))
    Dp = np.zeros((max(nw, 1), M, M))
    np.add.at(Bp, (np.searchsorted(pk, w_cw * nc + c_cw), r[mcw] % B, (c[mcw] - ncB) % M), v[mcw])
    np.add.at(Cp, (np.searchsorted(pk, w_wc * nc + c_wc), (r[mwc] - ncB) % M, c[mwc] % B), v[mwc])
    ww_r, ww_c = (r[mww] - ncB) // M, (c[mww] - ncB) // M
    if np.any(ww_r != ww_c):
        raise ValueError("well equations couple different wells")
    np.add.at(Dp, (ww_r, (r[mww] - ncB) % M, (c[mww] - ncB) % M), v[mww])
    return SplitSystem(rowptr, col, val, perf_ptr, perf_cells, Bp[:nperf], Cp[:nperf], Dp[:nw], dev, keys, pk)


DIRECT_MAX = 20000


def device_direct_solve(jac, rhs) -> np.ndarray:
    """jac x = rhs by a dense LU on the GPU (the reference's direct solve,
    on the device); up to DIRECT_MAX unknowns."""
    import torch
    n = jac.shape[0]
    if jac.shape[0] != jac.shape[1]:
        raise ValueError(f"a direct solve needs a square matrix, got {jac.shape}")
    if n > DIRECT_MAX:
        raise ValueError(f"a direct solve of {n} unknowns is beyond the device's dense limit ({DIRECT_MAX}); "
                         "give the model (.G, .n_phases, .fluid) for the CPR path")
    dev = torch.device("cuda")
    A = torch.as_tensor(sp.csr_array(jac).toarray(), dtype=torch.float64, device=dev)
    b = torch.as_tensor(np.asarray(rhs, dtype=float).ravel(), device=dev)
    return torch.linalg.solve(A, b).cpu().numpy()


def ruiz_equilibrate(a, sweeps: int = 3):
    """Row and column scalings dr, dc with dr A dc of unit max-norm rows and
    columns after `sweeps` alternating passes (the reference equilibrates
    the pressure system so, linear_solvers.py:205-224)."""
    a = sp.csr_array(a, dtype=float)
    dr, dc = np.ones(a.shape[0]), np.ones(a.shape[1])
    for _ in range(sweeps):
        b = sp.diags_array(dr) @ a @ sp.diags_array(dc)
        rmax = np.sqrt(abs(b).max(axis=1).toarray().ravel())
        rmax[rmax == 0.0] = 1.0
        dr = dr / rmax
        b = sp.diags_array(dr) @ a @ sp.diags_array(dc)
        cmax = np.sqrt(abs(b).max(axis=0).toarray().ravel())
        cmax[cmax == 0.0] = 1.0
        dc = dc / cmax
    return (sp.diags_array(dr) @ a @ sp.diags_array(dc)).tocsr(), dr, dc


def solve_reference_system(solver, jac, rhs, model, state) -> np.ndarray:
    """CPRSolver.solve's body: reorder, solve on the device, order back."""
    import warnings

    import torch
    missing = unmet_requirements(model)
    if missing:
        # the reference's fallback: loud when a model was given and falls
        # short, quiet when none was (the documented direct path)
        if model is not None:
            warnings.warn(f"CPR fell back to a direct solve: {'; '.join(missing)}. The answer is correct and "
                          "the cost is that of a full factorisation - set the missing attributes to get the "
                          "preconditioned path.", RuntimeWarning, stacklevel=3)
        solver.last_iterations = 0
        solver.last_fallback = tuple(missing)
        return device_direct_solve(jac, rhs)
    solver.last_fallback = ()
    if state is None:
        raise ValueError("the device CPR needs the state (its mean pressure sets the weights)")
    nc = int(model.G.cells.num)
    B = int(model.n_phases)
    if B not in (2, 3):
        raise ValueError(f"{B} phases: the device CPR handles two (oil-water) and three (black-oil)")
    sysd = getattr(solver, "_sys", None)
    fits = sysd is not None and sysd.B == B and sysd.nc == nc and sysd.nw == (jac.shape[0] - B * nc) // (B + 1)
    split = split_reference_jacobian(jac, nc, B, *((sysd.keys, sysd.pkeys) if fits else (None, None)))
    rowptr, col, val, perf_ptr, perf_cells, Bp, Cp, Dp, dev = split
    if not (fits and np.array_equal(split.keys, sysd.keys) and np.array_equal(split.pkeys, sysd.pkeys)):
        # a block or perforation the device system lacks: a new system on the
        # joined pattern (aggregation and lines are rebuilt)
        solver._sys = sysd = None
        sysd = DeviceSystem(B, nc, rowptr, col, perf_ptr, perf_cells)
        sysd.keys, sysd.pkeys = split.keys, split.pkeys
        solver._sys = sysd
    sysd.set_values(val, Bp, Cp, Dp)
    w = pressure_weights(model, state, B)
    elliptic = getattr(solver, "elliptic_solver", None)
    if elliptic is not None:
        # the caller's pressure solver on A_p = W J P, equilibrated as the
        # reference equilibrates it before handing it over
        a = sp.csr_array(jac)
        a_p = sum(w[i] * a[i * nc:(i + 1) * nc, :nc] for i in range(B))
        a_eq, dr, dc = ruiz_equilibrate(a_p)
        sysd.set_pressure_solver(lambda b: dc * np.asarray(elliptic.solve_linear_system(a_eq, dr * b)).ravel())
    rhs = np.asarray(rhs, dtype=float).ravel()
    rd = np.empty_like(rhs)
    rd[dev] = rhs
    try:
        x, rc, info = sysd.solve(solver.options(), w, torch.as_tensor(rd, device=sysd.device))
    finally:
        if elliptic is not None:
            sysd.set_pressure_solver(None)
    solver.last_iterations = int(info.iterations)
    solver.history.append((int(info.iterations), float(info.rel_residual)))
    if rc == 1:
        raise RuntimeError(f"CPR-FGMRES did not converge ({info.iterations} iterations, relative residual "
                           f"{info.rel_residual:.2e})")
    return x.cpu().numpy()[dev]


# -- the AD-OO seam (linear_solver_ad.py:55-140, 807-935) -----------------------
class LinearizedProblemAD:
    """A Newton system A dx = -b with its primary-variable layout: A, the
    residual b, var_names / var_sizes (the column blocks), state, dt; or
    built from AD equations (each with .val and .jac, one Jacobian block per
    primary variable) with from_equations."""

    def __init__(self, A=None, b=None, var_names=None, var_sizes=None, state=None, dt=None):
        self.A = A
        self.b = np.zeros(0) if b is None else np.asarray(b, dtype=float).ravel()
        self.var_names = None if var_names is None else list(var_names)
        self.var_sizes = None if var_sizes is None else [int(v) for v in var_sizes]
        self.state, self.dt = state, dt
        self.equations, self.types, self.equation_names = [], [], []
        self.num_vars = None

    @classmethod
    def from_equations(cls, equations, types, names, primary_vars, state=None, dt=0.0):
        prob = cls(state=state, dt=dt, var_names=[primary_vars] if isinstance(primary_vars, str) else primary_vars)
        prob.equations, prob.types, prob.equation_names = list(equations), list(types), list(names)
        return prob

    def assemble_system(self):
        """A and the residual b from the equations (combine_equations)."""
        if self.A is None and self.equations:
            blocks = [e for e in self.equations]
            self.b = np.concatenate([np.asarray(getattr(e, "val", e), dtype=float).ravel() for e in blocks])
            self.A = sp.vstack([sp.hstack(list(e.jac), format="csr") for e in blocks], format="csr")
        return self

    def num_eq(self) -> int:
        return len(self.equations)


def problem_system(problem):
    """(A, rhs) of a linearized problem, assembling it when it holds only its
    equations; rhs is minus the residual (a LinearizedProblemAD stores the
    residual, an MRST LinearizedProblem the right-hand side)."""
    assemble = getattr(problem, "assemble_system", None)
    if getattr(problem, "A", None) is None and callable(assemble):
        out = assemble()
        problem = out if out is not None else problem
    A, b = problem.A, np.asarray(problem.b, dtype=float).ravel()
    if A is None:
        raise ValueError("the linearized problem has neither a matrix nor equations")
    rhs = -b if type(problem).__name__ == "LinearizedProblemAD" else b
    return sp.csr_array(A), rhs


def split_increment(problem, result):
    """The increment as one block per primary variable (MRST's
    storeIncrements): widths from the problem's num_vars, else from its first
    equation's Jacobian blocks, else its var_sizes; whole otherwise."""
    result = np.asarray(result, dtype=float).ravel()
    eqs = [e for e in getattr(problem, "equations", []) if hasattr(e, "jac")]
    num_vars = getattr(problem, "num_vars", None)
    if eqs and num_vars is not None and sum(num_vars) == sum(j.shape[1] for j in eqs[0].jac):
        widths = list(num_vars)
    elif eqs:
        widths = [j.shape[1] for j in eqs[0].jac]
    elif getattr(problem, "var_sizes", None) and getattr(problem, "var_names", None):
        widths = list(problem.var_sizes)
    elif result.size == 0:
        return []
    else:
        return [result]
    edges = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    return [result[edges[i]:edges[i + 1]] for i in range(len(widths))]


def solve_linear_problem(solver, problem, model=None):
    """(dx, result, report) as the reference's solveLinearProblem returns:
    the device solve of A dx = -residual, split per primary variable, and a
    report with MRST's fields."""
    import time
    t0 = time.perf_counter()
    A, rhs = problem_system(problem)
    if not np.all(np.isfinite(rhs)):
        raise ValueError("Linear system rhs must have finite entries.")
    t_prep = time.perf_counter() - t0
    state = getattr(problem, "state", None)
    result = np.asarray(solver.solve(A, rhs, model=model, state=state, dt=getattr(problem, "dt", None)),
                        dtype=float).ravel()
    t_solve = time.perf_counter() - t0 - t_prep
    res = float(np.linalg.norm(A @ result - rhs) / max(np.linalg.norm(rhs), np.finfo(float).tiny))
    total = time.perf_counter() - t0
    report = {"Iterations": int(getattr(solver, "last_iterations", 0)), "Residual": res, "SolverTime": total,
              "LinearSolutionTime": t_solve, "PreparationTime": t_prep,
              "PostProcessTime": total - t_solve - t_prep, "Converged": True}
    return split_increment(problem, result), result, report
