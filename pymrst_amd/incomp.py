"""Incompressible TPFA pressure solve on the MI355X (incompTPFA), fed by a
device computeTrans.

Reference: incomp_tpfa (pymrst/modules/incomp/incomp_tpfa.py:55-211, the
full option surface) and incomp_tpfa_ad (:252-389, the variant MRST's
validation replays use), with compute_pressure_rhs
semantics (:392-435), incomp_well_pressure_drop
(pymrst/params/wells_and_bc/incomp_well_pressure_drop.py:9-26) and
compute_trans (pymrst/solvers/compute_trans.py:11-29).  The transmissibility,
assembly, AMG-preconditioned solve and flux reconstruction run on the GPU;
the host prepares boundary conditions and the wells' hydrostatic drops.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import NORM_GRAVITY, Grid, Rock, Well


# -- fluids, states, driving forces (pymrst/modules/incomp/fluid, solvers/init_res_sol.py) --
@dataclass
class IncompFluid:
    mu: np.ndarray
    rho: np.ndarray
    n: np.ndarray

    def mobility(self, s):
        s = np.atleast_2d(np.asarray(s, float))
        if self.mu.size == 1:
            kr = s[:, :1]
        else:
            s1 = s[:, 0]
            kr = np.column_stack([s1 ** self.n[0], (1.0 - s1) ** self.n[1]])
        return kr / self.mu


def init_single_fluid(mu: float, rho: float) -> IncompFluid:
    return IncompFluid(np.array([float(mu)]), np.array([float(rho)]), np.array([1.0]))


def init_simple_fluid(mu, rho, n=(1.0, 1.0)) -> IncompFluid:
    return IncompFluid(np.asarray(mu, float), np.asarray(rho, float), np.asarray(n, float))


@dataclass
class WellSol:
    flux: np.ndarray
    pressure: float
    cdp: np.ndarray = None


@dataclass
class State:
    pressure: np.ndarray
    flux: np.ndarray
    s: np.ndarray
    facePressure: np.ndarray = None
    wellSol: list = field(default_factory=list)
    A: object = None        # the system and its right-hand side, with MatrixOutput or reduce
    rhs: np.ndarray = None


def init_res_sol(g: Grid, p0, s0=1.0) -> State:
    nc = g.num_cells
    p = np.full(nc, float(p0)) if np.size(p0) == 1 else np.asarray(p0, float).copy()
    s = np.atleast_2d(np.asarray(s0, float))
    s = np.repeat(s, nc, axis=0) if s.shape[0] == 1 else s
    return State(pressure=p, flux=np.zeros(g.num_faces), s=s)


@dataclass
class BoundaryCondition:
    face: np.ndarray
    type: List[str]
    value: np.ndarray


def add_bc(bc: Optional[BoundaryCondition], faces, kind: str, value) -> BoundaryCondition:
    faces = np.asarray(faces, dtype=np.int64).ravel()
    vals = np.broadcast_to(np.asarray(value, float), faces.shape).copy()
    if kind not in ("pressure", "flux"):
        raise ValueError(f"unsupported boundary condition '{kind}'")
    if bc is None:
        return BoundaryCondition(faces, [kind] * faces.size, vals)
    return BoundaryCondition(np.concatenate([bc.face, faces]), bc.type + [kind] * faces.size,
                             np.concatenate([bc.value, vals]))


@dataclass
class Source:
    cell: np.ndarray
    rate: np.ndarray


def add_source(src: Optional[Source], cells, rate) -> Source:
    cells = np.asarray(cells, dtype=np.int64).ravel()
    rate = np.broadcast_to(np.asarray(rate, float), cells.shape).copy()
    if src is None:
        return Source(cells, rate)
    return Source(np.concatenate([src.cell, cells]), np.concatenate([src.rate, rate]))


def well_pressure_drop(w: Well, mob: np.ndarray, rho: np.ndarray, norm_g: float) -> np.ndarray:
    """incomp_well_pressure_drop.py:9-26: injector mixture from compi,
    producer from WI-weighted fractional-flow densities."""
    nph = rho.size
    if w.sign > 0 or (w.type == "rate" and w.val > 0):
        n = min(nph, w.compi.size)
        rho_mix = float(rho[:n] @ w.compi[:n])
    else:
        mw = mob[w.cells]
        frac = mw / mw.sum(axis=1, keepdims=True)
        n = min(nph, mob.shape[1])
        rho_mix = float((w.WI * (frac[:, :n] @ rho[:n])).sum() / w.WI.sum())
    return norm_g * w.dZ * rho_mix


# -- device solver -----------------------------------------------------------------
class IncompTPFA:
    """Device handle for one grid: computeTrans and pressure solves."""

    def __init__(self, g: Grid, device=None, rtol: float = 1e-12, maxiter: int = 500):
        if not torch.cuda.is_available():
            raise _lib.PmxError("incompTPFA on the device needs a GPU (no CPU fallback)")
        self.lib = _lib.load()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.g = g
        self.rtol, self.maxiter = rtol, maxiter
        nc, nf = g.num_cells, g.num_faces
        self.face_pos = np.ascontiguousarray(np.arange(nc + 1) * 6, dtype=np.int32)
        self.hf_face = np.ascontiguousarray(g.cell_faces.ravel(), dtype=np.int32)
        nb = np.ascontiguousarray(g.neighbors.ravel(), dtype=np.int32)
        cc = np.ascontiguousarray(g.cell_centroids, dtype=np.float64)
        fc = np.ascontiguousarray(g.face_centroids, dtype=np.float64)
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            h = ctypes.c_void_p()
            _lib.check(self.lib.pmx_incomp_create(self.device.index or 0, ctypes.c_void_p(self.stream.cuda_stream),
                                                  ctypes.byref(h)), "pmx_incomp_create")
            self.h = h
            _lib.check(self.lib.pmx_incomp_set_grid(h, nc, nf, _lib.ptr(self.face_pos), _lib.ptr(self.hf_face),
                                                    _lib.ptr(nb), _lib.ptr(cc), _lib.ptr(fc)), "pmx_incomp_set_grid")
        self._normals = torch.as_tensor(g.face_normals, dtype=torch.float64, device=self.device).contiguous()

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and getattr(self, "lib", None) is not None:
            try:
                self.lib.pmx_incomp_destroy(h)
            except Exception:
                pass
            self.h = None

    def compute_trans(self, rock: Rock) -> torch.Tensor:
        """Half-face transmissibilities on the device (cell-major W,E,S,N,T,B)."""
        perm = torch.as_tensor(np.ascontiguousarray(rock.perm, dtype=np.float64), device=self.device)
        T = torch.empty(self.hf_face.size, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.pmx_compute_trans(self.h, _lib.ptr(self._normals), int(rock.perm.shape[1]),
                                              _lib.ptr(perm), _lib.ptr(T)), "pmx_compute_trans")
        return T

    def solve(self, state: State, T, fluid: IncompFluid, bc=None, src=None, wells=None, gravity=None, *,
              bcp=None, lin_solve=None, matrix_output=False, reduce=False, use_trans=False,
              condition_number=False, well_gravity=None) -> State:
        """Assemble on the device, solve (on the device, or by `lin_solve(A,
        b)` on the exported system), reconstruct on the device.
        `well_gravity`: the gravity magnitude of the wells' hydrostatic
        drops (default: that of `gravity`)."""
        import scipy.sparse as sp
        g = self.g
        nc, nf = g.num_cells, g.num_faces
        wells = list(wells or [])
        gvec = np.zeros(3) if gravity is None else np.resize(np.asarray(gravity, float), 3)
        mob = fluid.mobility(state.s)
        totmob = mob.sum(axis=1)
        omega = (mob * fluid.rho).sum(axis=1) / totmob
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)
        if use_trans:
            T_d = dev(self._faces_to_half_faces(np.asarray(T.cpu() if isinstance(T, torch.Tensor) else T, float),
                                                totmob))
        else:
            T_d = T if isinstance(T, torch.Tensor) else dev(T)
        mob_d, om_d = dev(totmob), dev(omega)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        if bc is not None and bc.face.size:
            bc_face, bc_type, bc_val = i32(bc.face), i32([0 if t == "pressure" else 1 for t in bc.type]), f64(bc.value)
            nbc = bc_face.size
        else:
            bc_face = bc_type = bc_val = None
            nbc = 0
        if src is not None and src.cell.size:
            s_cell, s_rate, nsrc = i32(src.cell), f64(src.rate), src.cell.size
        else:
            s_cell = s_rate = None
            nsrc = 0
        if bcp is not None and np.size(bcp.face):
            j_face, j_val, njump = i32(bcp.face), f64(np.broadcast_to(bcp.value, np.shape(bcp.face))), np.size(bcp.face)
        else:
            j_face = j_val = None
            njump = 0
        nw = len(wells)
        norm_g = float(np.linalg.norm(gvec))
        wg = norm_g if well_gravity is None else float(well_gravity)
        cdp = [well_pressure_drop(w, mob, fluid.rho, wg) for w in wells]
        if nw:
            perf_ptr = i32(np.concatenate([[0], np.cumsum([w.cells.size for w in wells])]))
            perf_cells = i32(np.concatenate([w.cells for w in wells]))
            WI = f64(np.concatenate([w.WI for w in wells]))
            wtype = i32([0 if w.type == "bhp" else 1 for w in wells])
            wval = f64([w.val for w in wells])
            wdp = f64(np.concatenate(cdp))
            well_p = np.zeros(nw)
            well_flux = np.zeros(perf_cells.size)
        else:
            perf_ptr = perf_cells = WI = wtype = wval = wdp = well_p = well_flux = None
        P = _lib.ptr
        _lib.check(self.lib.pmx_incomp_assemble(self.h, P(T_d), P(mob_d), P(om_d),
                                                P(f64(gvec)) if norm_g > 0 else P(None), nbc, P(bc_face),
                                                P(bc_type), P(bc_val), nsrc, P(s_cell), P(s_rate), njump,
                                                P(j_face), P(j_val), nw, P(perf_ptr), P(perf_cells), P(WI),
                                                P(wtype), P(wval), P(wdp)), "pmx_incomp_assemble")
        A = rhs = None
        if matrix_output or lin_solve is not None or reduce or condition_number:
            A, rhs = self._export(nw, perf_ptr, perf_cells)
            if condition_number:
                print(f"Condition number is: {np.linalg.cond(A.toarray())}")
        x = torch.empty(nc + nw, dtype=torch.float64, device=self.device)
        A_out, rhs_out = A, rhs
        if reduce and nw:
            # Schur complement of the well unknowns (incomp_tpfa.py:173-182)
            C, Dw = A[nc:, :nc], A[nc:, nc:].diagonal()
            Dinv = sp.diags_array(1.0 / Dw)
            A_out = (A[:nc, :nc] - C.T @ Dinv @ C).tocsr()
            rhs_out = rhs[:nc] - C.T @ (Dinv @ rhs[nc:])
        if lin_solve is not None:
            xs = np.asarray(lin_solve(A_out, rhs_out), dtype=float).ravel()
            if reduce and nw:
                xs = np.concatenate([xs, (rhs[nc:] - A[nc:, :nc] @ xs) / A[nc:, nc:].diagonal()])
            x.copy_(torch.as_tensor(xs, device=self.device))
            self.last_info = None
        else:
            o = _lib.solver_defaults()
            o.rtol, o.maxiter, o.restart = self.rtol, self.maxiter, 100
            info = _lib.SolveInfo()
            rc = self.lib.pmx_incomp_solve_assembled(self.h, ctypes.byref(o), P(x), ctypes.byref(info))
            _lib.check(rc, "pmx_incomp_solve_assembled")
            if rc != 0:
                raise RuntimeError(f"incompTPFA pressure solve did not converge ({info.iterations} its, "
                                   f"rel residual {info.rel_residual:.1e})")
            self.last_info = info
        p = torch.empty(nc, dtype=torch.float64, device=self.device)
        flux = torch.empty(nf, dtype=torch.float64, device=self.device)
        fpress = torch.empty(nf, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.pmx_incomp_reconstruct(self.h, P(x), P(p), P(flux), P(fpress), P(well_p), P(well_flux)),
                   "pmx_incomp_reconstruct")
        out = State(pressure=p.cpu().numpy(), flux=flux.cpu().numpy(), s=state.s,
                    facePressure=fpress.cpu().numpy())
        if matrix_output or reduce:
            out.A, out.rhs = A_out, rhs_out
        off = 0
        for k, w in enumerate(wells):
            n = w.cells.size
            out.wellSol.append(WellSol(flux=well_flux[off:off + n].copy(), pressure=float(well_p[k]), cdp=cdp[k]))
            off += n
        return out

    def _export(self, nw, perf_ptr, perf_cells):
        """The last assembled system as a SciPy CSR matrix over [cells, wells]
        and its right-hand side."""
        import scipy.sparse as sp
        nc = self.g.num_cells
        nnz = ctypes.c_int32()
        _lib.check(self.lib.pmx_incomp_nnz(self.h, ctypes.byref(nnz)), "pmx_incomp_nnz")
        rowptr = np.zeros(nc + 1, np.int32)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        npf = 0 if not nw else int(perf_ptr[-1])
        Bp, Cp, D = np.zeros(max(npf, 1)), np.zeros(max(npf, 1)), np.zeros(max(nw, 1))
        rhs = np.zeros(nc + nw)
        P = _lib.ptr
        _lib.check(self.lib.pmx_incomp_export(self.h, P(rowptr), P(col), P(val), P(Bp), P(Cp), P(D), P(rhs)),
                   "pmx_incomp_export")
        rows = np.repeat(np.arange(nc), np.diff(rowptr))
        if nw:
            wrow = np.repeat(np.arange(nw), np.diff(perf_ptr)) + nc
            r = np.concatenate([rows, perf_cells, wrow, nc + np.arange(nw)])
            c = np.concatenate([col, wrow, perf_cells, nc + np.arange(nw)])
            v = np.concatenate([val, Bp[:npf], Cp[:npf], D[:nw]])
        else:
            r, c, v = rows, col, val
        A = sp.csr_array((v, (r, c)), shape=(nc + nw, nc + nw))
        A.sum_duplicates()
        A.eliminate_zeros()
        return A, rhs

    def _faces_to_half_faces(self, T_face, totmob):
        """One transmissibility per interface (use_trans) as the one-sided
        values the device assembly takes: the reference's face value
        T_f x mean(mobility of the face's cells) (incomp_tpfa.py:223-236) is
        what the device's harmonic mean returns when each side carries
        T_f x mean x (cells on the face) / own mobility."""
        g = self.g
        T_face = np.asarray(T_face, float).ravel()
        if T_face.size != g.num_faces:
            raise ValueError(f"use_trans: one transmissibility per interface ({g.num_faces}) expected, "
                             f"got {T_face.size}")
        hf_face = g.cell_faces.ravel()
        hf_cell = np.repeat(np.arange(g.num_cells), g.cell_faces.shape[1])
        count = (np.asarray(g.neighbors) >= 0).sum(axis=1)
        fmob = np.bincount(hf_face, totmob[hf_cell], minlength=g.num_faces) / count
        ft = T_face * fmob
        return ft[hf_face] * count[hf_face] / totmob[hf_cell]


_handles = {}


def _handle(g: Grid, device=None) -> IncompTPFA:
    key = (id(g), str(device))
    h = _handles.get(key)
    if h is None or h.g is not g:
        h = IncompTPFA(g, device)
        _handles[key] = h
    return h


def compute_trans(g: Grid, rock: Rock, device=None) -> np.ndarray:
    """computeTrans on the device; returns the half-face array on the host."""
    return _handle(g, device).compute_trans(rock).cpu().numpy()


def incomp_tpfa_ad(state: State, g: Grid, T, fluid: IncompFluid, bc=None, src=None, wells=None, gravity=None,
                   device=None) -> State:
    """incomp_tpfa_ad (incomp_tpfa.py:252-389): a new state with pressure,
    flux, facePressure and wellSol; `gravity` None means none."""
    return _handle(g, device).solve(state, T, fluid, bc=bc, src=src, wells=wells, gravity=gravity)


def incomp_tpfa(state: State, G: Grid, T, fluid: IncompFluid, bc=None, src=None, wells=None, gravity=None, *,
                W=None, bcp=None, LinSolve=None, MatrixOutput: bool = False, Verbose: bool = False,
                condition_number: bool = False, pc_form: str = "nonwetting", reduce: bool = False,
                use_trans: bool = False, device=None) -> State:
    """incomp_tpfa (incomp_tpfa.py:55-211): a new state with pressure, flux,
    facePressure and wellSol, and with MatrixOutput (or reduce) the system
    A, rhs (the wells' Schur complement with reduce).  `gravity` defaults to
    the process-wide setting (:mod:`pymrst_amd.gravity`), which also sets
    the wells' hydrostatic drops, as in the reference.  `W` is the legacy
    name of `wells`; `bcp` adds periodic pressure jumps across its faces;
    `LinSolve(A, b) -> x` replaces the device solve; `use_trans` takes one
    transmissibility per interface.  Assembly and reconstruction run on the
    device in every case."""
    from .gravity import current_gravity
    if W is not None:
        if wells is not None:
            raise ValueError("give the wells as `wells` or as `W`, not both")
        wells = W
    if getattr(fluid, "pc", None) is not None:
        raise NotImplementedError("capillary pressure is not supported by the device incompTPFA")
    gvec = current_gravity() if gravity is None else np.resize(np.asarray(gravity, float), 3)
    if not MatrixOutput and bc is None and src is None and bcp is None and not wells and not np.any(gvec):
        import warnings
        warnings.warn("No external driving forces present in model--state remains unchanged.")
    out = _handle(G, device).solve(state, T, fluid, bc=bc, src=src, wells=wells, gravity=gvec, bcp=bcp,
                                   lin_solve=LinSolve, matrix_output=MatrixOutput, reduce=reduce,
                                   use_trans=use_trans, condition_number=condition_number,
                                   well_gravity=float(np.linalg.norm(current_gravity())))
    if Verbose:
        print(f"incompTPFA: {G.num_cells} cells, {len(wells or [])} wells")
    return out


incompTPFA = incomp_tpfa


def gravity_vector(dim: int = 3) -> np.ndarray:
    v = np.zeros(dim)
    v[-1] = NORM_GRAVITY
    return v
