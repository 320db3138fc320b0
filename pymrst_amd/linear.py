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
Unlike the reference, a model that lacks what the pressure stage needs is
an error here, not a direct solve on the host.
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

    def solve(self, opts, w, rhs):
        import torch
        x = torch.zeros_like(rhs)
        info = _lib.SolveInfo()
        wv = np.ascontiguousarray(w, dtype=np.float64)
        rc = self.lib.pmx_sys_solve(self.h, ctypes.byref(opts), _lib.ptr(wv), _lib.ptr(rhs), _lib.ptr(x),
                                    ctypes.byref(info))
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


def solve_reference_system(solver, jac, rhs, model, state) -> np.ndarray:
    """CPRSolver.solve's body: reorder, solve on the device, order back."""
    import torch
    missing = unmet_requirements(model)
    if missing:
        raise ValueError("the device CPR needs " + "; ".join(missing) + " (there is no host direct fallback)")
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
    rhs = np.asarray(rhs, dtype=float).ravel()
    rd = np.empty_like(rhs)
    rd[dev] = rhs
    x, rc, info = sysd.solve(solver.options(), w, torch.as_tensor(rd, device=sysd.device))
    solver.last_iterations = int(info.iterations)
    solver.history.append((int(info.iterations), float(info.rel_residual)))
    if rc == 1:
        raise RuntimeError(f"CPR-FGMRES did not converge ({info.iterations} iterations, relative residual "
                           f"{info.rel_residual:.2e})")
    return x.cpu().numpy()[dev]
